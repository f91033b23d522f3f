"""The oracle: CPU restatements of the reference's projection path.

Test infrastructure only.  ``ref_numpy`` is the independent numpy
restatement; ``ref_c`` loads the C restatement (``lib/libdfp_oracle.so``,
built by ``make -C oracle``) that the GPU parity tests and the host-CPU
baseline use.  The product never imports this package.
"""
