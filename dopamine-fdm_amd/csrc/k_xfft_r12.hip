// x-direction spectral kernels with a radix-12 first z stage
#include "k_xdct_impl.h"

DFP_INSTANTIATE_XFFT(12)
DFP_INSTANTIATE_XDCT(12)
