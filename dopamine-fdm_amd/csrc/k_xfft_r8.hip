// x-direction spectral kernels with a radix-8 first z stage
#include "k_xdct_impl.h"

DFP_INSTANTIATE_XFFT(8)
DFP_INSTANTIATE_XDCT(8)
