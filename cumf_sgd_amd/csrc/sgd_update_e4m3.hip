// Batch-Hogwild update instantiations for e4m3 factor storage.
#include "sgd_update.h"

namespace cumf {

template <>
hipError_t launch_sgd_prec<kE4M3>(const SgdLaunch& a, int blocks, int k, int G, bool coh, int D, hipStream_t s)
{
    switch (k) {
    case 32: return launch_sgd_g<32, kE4M3>(a, blocks, G, coh, D, s);
    case 64: return launch_sgd_g<64, kE4M3>(a, blocks, G, coh, D, s);
    case 128: return launch_sgd_g<128, kE4M3>(a, blocks, G, coh, D, s);
    }
    return hipErrorInvalidValue;
}

template <>
hipError_t launch_batched_for<kE4M3>(const SgdLaunch& a, int blocks, int D, hipStream_t s)
{
    return launch_batched_prec<kE4M3>(a, blocks, D, s);
}

template <>
hipError_t launch_tile_prec<kE4M3>(const SgdLaunch& a, int blocks, int k, int G, int D, int waves, size_t lds, hipStream_t s)
{
    switch (k) {
    case 32: return launch_tile_k<32, kE4M3>(a, blocks, G, D, waves, lds, s);
    case 64: return launch_tile_k<64, kE4M3>(a, blocks, G, D, waves, lds, s);
    case 128: return launch_tile_k<128, kE4M3>(a, blocks, G, D, waves, lds, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace cumf
