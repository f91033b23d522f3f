// x-direction spectral kernels with a radix-1 first z stage
#include "k_xfft_impl.h"

DFP_INSTANTIATE_XFFT(1)
