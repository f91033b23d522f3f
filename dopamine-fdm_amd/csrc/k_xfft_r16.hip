// x-direction spectral kernels with a radix-16 first z stage
#include "k_xdct_impl.h"

DFP_INSTANTIATE_XFFT(16)
DFP_INSTANTIATE_XDCT(16)
