// Batch-Hogwild update instantiations for bf16 factor storage with stochastic rounding.
#include "sgd_update.h"

namespace cumf {

template <>
hipError_t launch_sgd_prec<kBF16SR>(const SgdLaunch& a, int blocks, int k, int G, bool coh, int D, hipStream_t s)
{
    switch (k) {
    case 32: return launch_sgd_g<32, kBF16SR>(a, blocks, G, coh, D, s);
    case 64: return launch_sgd_g<64, kBF16SR>(a, blocks, G, coh, D, s);
    case 128: return launch_sgd_g<128, kBF16SR>(a, blocks, G, coh, D, s);
    }
    return hipErrorInvalidValue;
}

template <>
hipError_t launch_batched_for<kBF16SR>(const SgdLaunch& a, int blocks, int D, hipStream_t s)
{
    return launch_batched_prec<kBF16SR>(a, blocks, D, s);
}

template <>
hipError_t launch_tile_prec<kBF16SR>(const SgdLaunch& a, int blocks, int k, int G, int D, int waves, size_t lds, hipStream_t s)
{
    switch (k) {
    case 32: return launch_tile_k<32, kBF16SR>(a, blocks, G, D, waves, lds, s);
    case 64: return launch_tile_k<64, kBF16SR>(a, blocks, G, D, waves, lds, s);
    case 128: return launch_tile_k<128, kBF16SR>(a, blocks, G, D, waves, lds, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace cumf
