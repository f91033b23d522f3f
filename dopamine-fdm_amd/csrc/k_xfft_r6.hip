// x-direction spectral kernels with a radix-6 first z stage
#include "k_xdct_impl.h"

DFP_INSTANTIATE_XFFT(6)
DFP_INSTANTIATE_XDCT(6)
