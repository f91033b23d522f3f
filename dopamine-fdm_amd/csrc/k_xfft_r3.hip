// x-direction spectral kernels with a radix-3 first z stage
#include "k_xdct_impl.h"

DFP_INSTANTIATE_XFFT(3)
DFP_INSTANTIATE_XDCT(3)
