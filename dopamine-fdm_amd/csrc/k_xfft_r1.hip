// x-direction spectral kernels with a radix-1 first z stage
#include "k_xdct_impl.h"

DFP_INSTANTIATE_XFFT(1)
DFP_INSTANTIATE_XDCT(1)
