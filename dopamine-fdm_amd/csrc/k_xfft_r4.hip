// x-direction spectral kernels with a radix-4 first z stage
#include "k_xdct_impl.h"

DFP_INSTANTIATE_XFFT(4)
DFP_INSTANTIATE_XDCT(4)
