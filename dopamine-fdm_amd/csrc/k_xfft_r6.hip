// x-direction spectral kernels with a radix-6 first z stage
#include "k_xfft_impl.h"

DFP_INSTANTIATE_XFFT(6)
