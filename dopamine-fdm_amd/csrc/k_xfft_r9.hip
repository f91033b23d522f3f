// x-direction spectral kernels with a radix-9 first z stage
#include "k_xdct_impl.h"

DFP_INSTANTIATE_XFFT(9)
DFP_INSTANTIATE_XDCT(9)
