// kmpm_tile_launch.h -- launchers of the tile kernels for one (dimension,
// shape function, tile variant).  Included only by kmpm_tiles.hip, which
// instantiates TileLaunch for one variant per object file.
#pragma once

#include "kmpm_kernels.h"

namespace kmpm {

// One P2G launch: the step's P2G or the MUSL re-map (momentum only, which
// has no force term: its APIC form is XF 1 whatever the force).
template <int DIM, int S1, int TX, int TY, int TZ, int PREC, bool TL, int XF>
static void p2g_go(kmpm_ctx *c, int g, int remap) {
  if (remap)
    hipLaunchKernelGGL((k_p2g<DIM, S1, TX, TY, TZ, true, PREC, TL, XF == 2 ? 1 : XF>), dim3(g),
                       dim3(256), 0, c->stream, c->a);
  else
    hipLaunchKernelGGL((k_p2g<DIM, S1, TX, TY, TZ, false, PREC, TL, XF>), dim3(g), dim3(256), 0,
                       c->stream, c->a);
}

template <int DIM, int S1, int V>
hipError_t TileLaunch<DIM, S1, V>::p2g(kmpm_ctx *c, int remap) {
  using TT = Tiles<DIM, S1, V>;
  if (c->a.np == 0) return hipSuccess;
  const int g = persistent_grid(c, 8);
  const bool tl = c->a.tl != 0;
  const int xf = c->a.xf;
  if constexpr (V != 0) {
    if (xf) return hipErrorInvalidValue;   // the APIC kernels exist for variant 0 only
  }
  PREC_DISPATCH(c, {
    if constexpr (V == 0) {
      if (xf && tl) p2g_go<DIM, S1, TT::TX, TT::TY, TT::TZ, PREC, true, 1>(c, g, remap);
      else if (xf == 2) p2g_go<DIM, S1, TT::TX, TT::TY, TT::TZ, PREC, false, 2>(c, g, remap);
      else if (xf) p2g_go<DIM, S1, TT::TX, TT::TY, TT::TZ, PREC, false, 1>(c, g, remap);
    }
    if (!xf) {
      if (tl) p2g_go<DIM, S1, TT::TX, TT::TY, TT::TZ, PREC, true, 0>(c, g, remap);
      else p2g_go<DIM, S1, TT::TX, TT::TY, TT::TZ, PREC, false, 0>(c, g, remap);
    }
  });
  return hipGetLastError();
}

template <int DIM, int S1, int V>
hipError_t TileLaunch<DIM, S1, V>::grid_update(kmpm_ctx *c, int mode) {
  using TT = Tiles<DIM, S1, V>;
  int64_t total = (int64_t)c->a.tg.ntiles * c->a.tg.T[0] * c->a.tg.T[1] * c->a.tg.T[2];
  int g = nblk(total, 256);
  if (g > c->cu_count * 8) g = c->cu_count * 8;
  if (g < 1) g = 1;
  PREC_DISPATCH(c, {
    hipLaunchKernelGGL((k_grid_update<DIM, S1, TT::TX, TT::TY, TT::TZ, PREC>), dim3(g), dim3(256),
                       0, c->stream, c->a, mode);
  });
  return hipGetLastError();
}

template <int DIM, int S1, int V>
hipError_t TileLaunch<DIM, S1, V>::grid_clean(kmpm_ctx *c) {
  using TT = Tiles<DIM, S1, V>;
  int g = nblk(c->a.nn, 256);
  if (g > c->cu_count * 16) g = c->cu_count * 16;
  PREC_DISPATCH(c, {
    hipLaunchKernelGGL((k_grid_clean<DIM, TT::TX, TT::TY, TT::TZ, PREC>), dim3(g), dim3(256), 0,
                       c->stream, c->a);
  });
  return hipGetLastError();
}

template <int DIM, int S1, int V>
hipError_t TileLaunch<DIM, S1, V>::g2p_usl(kmpm_ctx *c) {
  using TT = Tiles<DIM, S1, V>;
  if (c->a.np == 0) return hipSuccess;
  const int g = persistent_grid(c, 8);
  if (c->a.tl) {
    MC_DISPATCH(c, PREC_DISPATCH(c, {
      hipLaunchKernelGGL((k_g2p_usl_tl<DIM, S1, TT::TX, TT::TY, TT::TZ, MC, PREC>), dim3(g),
                         dim3(256), 0, c->stream, c->a);
    }));
    return hipGetLastError();
  }
  // a learned constitutive update (the surrogate's networks or the cut
  // network) runs in its own kernel after this one, which stages its inputs
  const bool learned = learned_constitutive(c);
  if constexpr (V != 0) {
    if (c->a.xf) return hipErrorInvalidValue;   // the APIC kernels exist for variant 0 only
  } else if (c->a.xf) {
    MC_DISPATCH(c, PREC_DISPATCH(c, {
      hipLaunchKernelGGL((k_g2p_usl<DIM, S1, TT::TX, TT::TY, TT::TZ, MC, PREC, false, true>),
                         dim3(g), dim3(256), 0, c->stream, c->a);
    }));
    return hipGetLastError();
  }
  MC_DISPATCH(c, PREC_DISPATCH(c, {
    constexpr bool CAN_SURR = MC == MC_JC_SHOCK && PREC == KMPM_PRECISION_MIXED;
    if (CAN_SURR && learned)
      hipLaunchKernelGGL((k_g2p_usl<DIM, S1, TT::TX, TT::TY, TT::TZ, MC, PREC, CAN_SURR>),
                         dim3(g), dim3(256), 0, c->stream, c->a);
    else
      hipLaunchKernelGGL((k_g2p_usl<DIM, S1, TT::TX, TT::TY, TT::TZ, MC, PREC, false>), dim3(g),
                         dim3(256), 0, c->stream, c->a);
  }));
  return hipGetLastError();
}

template <int DIM, int S1, int V>
hipError_t TileLaunch<DIM, S1, V>::g2p_musl_a(kmpm_ctx *c) {
  using TT = Tiles<DIM, S1, V>;
  if (c->a.np == 0) return hipSuccess;
  const int g = persistent_grid(c, 8);
  if (c->a.tl) {
    PREC_DISPATCH(c, {
      hipLaunchKernelGGL((k_g2p_musl_a<DIM, S1, TT::TX, TT::TY, TT::TZ, PREC, true>), dim3(g),
                         dim3(256), 0, c->stream, c->a);
    });
    return hipGetLastError();
  }
  PREC_DISPATCH(c, {
    hipLaunchKernelGGL((k_g2p_musl_a<DIM, S1, TT::TX, TT::TY, TT::TZ, PREC, false>), dim3(g),
                       dim3(256), 0, c->stream, c->a);
  });
  return hipGetLastError();
}

template <int DIM, int S1, int V>
hipError_t TileLaunch<DIM, S1, V>::stress_musl_b(kmpm_ctx *c) {
  using TT = Tiles<DIM, S1, V>;
  if (c->a.np == 0) return hipSuccess;
  const int g = persistent_grid(c, 8);
  if (c->a.tl) {
    MC_DISPATCH(c, PREC_DISPATCH(c, {
      hipLaunchKernelGGL((k_stress_musl_b_tl<DIM, S1, TT::TX, TT::TY, TT::TZ, MC, PREC>), dim3(g),
                         dim3(256), 0, c->stream, c->a);
    }));
    return hipGetLastError();
  }
  MC_DISPATCH(c, PREC_DISPATCH(c, {
    hipLaunchKernelGGL((k_stress_musl_b<DIM, S1, TT::TX, TT::TY, TT::TZ, MC, PREC>), dim3(g),
                       dim3(256), 0, c->stream, c->a);
  }));
  return hipGetLastError();
}

// The thermal transfer and grid temperature (thermo-mechanical runs: FP64,
// tile variant 0).
template <int DIM, int S1, int V>
hipError_t TileLaunch<DIM, S1, V>::p2g_thermal(kmpm_ctx *c, int remap) {
  using TT = Tiles<DIM, S1, V>;
  if constexpr (V != 0) {
    return hipErrorInvalidValue;
  } else {
    if (c->a.np == 0) return hipSuccess;
    const int g = persistent_grid(c, 8);
    constexpr int FP = KMPM_PRECISION_FP64;
    if (c->a.tl) {
      if (remap)
        hipLaunchKernelGGL((k_p2g<DIM, S1, TT::TX, TT::TY, TT::TZ, true, FP, true, 0, true>),
                           dim3(g), dim3(256), 0, c->stream, c->a);
      else
        hipLaunchKernelGGL((k_p2g<DIM, S1, TT::TX, TT::TY, TT::TZ, false, FP, true, 0, true>),
                           dim3(g), dim3(256), 0, c->stream, c->a);
    } else {
      if (remap)
        hipLaunchKernelGGL((k_p2g<DIM, S1, TT::TX, TT::TY, TT::TZ, true, FP, false, 0, true>),
                           dim3(g), dim3(256), 0, c->stream, c->a);
      else
        hipLaunchKernelGGL((k_p2g<DIM, S1, TT::TX, TT::TY, TT::TZ, false, FP, false, 0, true>),
                           dim3(g), dim3(256), 0, c->stream, c->a);
    }
    return hipGetLastError();
  }
}

template <int DIM, int S1, int V>
hipError_t TileLaunch<DIM, S1, V>::grid_thermal(kmpm_ctx *c, int mode) {
  using TT = Tiles<DIM, S1, V>;
  if constexpr (V != 0) {
    return hipErrorInvalidValue;
  } else {
    int64_t total = (int64_t)c->a.tg.ntiles * c->a.tg.T[0] * c->a.tg.T[1] * c->a.tg.T[2];
    int g = nblk(total, 256);
    if (g > c->cu_count * 8) g = c->cu_count * 8;
    if (g < 1) g = 1;
    hipLaunchKernelGGL((k_grid_thermal<DIM, S1, TT::TX, TT::TY, TT::TZ>), dim3(g), dim3(256), 0,
                       c->stream, c->a, mode);
    return hipGetLastError();
  }
}

}  // namespace kmpm
