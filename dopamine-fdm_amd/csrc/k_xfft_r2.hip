// x-direction spectral kernels with a radix-2 first z stage
#include "k_xfft_impl.h"

DFP_INSTANTIATE_XFFT(2)
