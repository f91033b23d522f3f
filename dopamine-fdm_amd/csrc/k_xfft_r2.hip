// x-direction spectral kernels with a radix-2 first z stage
#include "k_xdct_impl.h"

DFP_INSTANTIATE_XFFT(2)
DFP_INSTANTIATE_XDCT(2)
