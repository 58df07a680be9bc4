// wg_kernels.hip -- HIP kernels for gfx950: the daily vertical water balance and
// the lateral routing of WaterGAP 2 (parity option profile).
//
// Two kernels per simulated day:
//   wg_cell_kernel   one lane per (member, cell).  It finishes the previous day
//                    (reduction factors, river width, surface-water and land area
//                    fractions: routing.cpp:2478-2488, 2792-2802, 3071-3078,
//                    3286-3295, 3546-3586, 5034-5188, 5343-5352), runs calcNewDay
//                    (daily.cpp:94-1279, lai.cpp:152-306) and the part of the
//                    routing loop that reads no upstream water (routing.cpp:1877-2617),
//                    and for cells without upstream cells the whole rest of the
//                    routing loop as well.
//   wg_chain_kernel  one workgroup per (routing group, member).  It walks the Kahn
//                    levels of the group with workgroup barriers; each cell gathers
//                    the transported volume of its upstream cells in routing order
//                    (the order of the serial += at routing.cpp:3955-3958) and runs
//                    routing.cpp:2623-3958 (global lake, reservoir, global wetland,
//                    arid groundwater, river).
// Arithmetic follows the reference expression by expression (operand order and
// association), compiled with -ffp-contract=off, in FP64 with the reference's
// float/short inputs promoted exactly where the reference promotes them; exp and
// pow are the deterministic ones of wg_libm.h, shared with the CPU oracle.
#include <hip/hip_runtime.h>

#include "wg_device.h"
// The exp/log/pow table lives in LDS, one copy per wavefront (up to 4 per
// workgroup): a lane-varying index into a global table costs a vector-memory
// round trip per lookup, two per pow.
__shared__ double wg_exp2_lds[4][128];
#define WG_EXP2_TAB_LDS (wg_exp2_lds[threadIdx.x >> 6])
#include "wg_libm.h"

namespace {
// Copies the table into the wavefront's LDS copy with one direct global-to-LDS load
// (16 bytes per lane, lane i to bytes 16i..16i+15).  The compiler waits for it
// before the first LDS read; workgroups of at most 256 threads.
__device__ inline void load_exp2_table() {
  const int lane = threadIdx.x & 63;
  __builtin_amdgcn_global_load_lds((const void *)(wg_exp2_tab_d + 2 * lane),
                                   (__attribute__((address_space(3))) void *)&wg_exp2_lds[threadIdx.x >> 6][0], 16,
                                   0, 0);
}
}  // namespace

namespace {

#include "wg_route_step.h"

// ---- ensemble perturbation: Philox4x32-10 + Box-Muller normal deviates ----
__device__ inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Two independent standard normal deviates for (cell, day) of one member: one
// Philox block gives u1, u2 = (c + 1/2) 2^-32 in (0, 1); Box-Muller
// R = sqrt(-2 ln u1), z = (R cos 2 pi u2, R sin 2 pi u2).  ln, sin and cos are the
// deterministic ones shared with the host, sqrt is correctly rounded, so the
// oracle draws the same bits.  |z| < 6.7 (u1 >= 2^-33).
__device__ inline void normal_pair(uint64_t seed, uint32_t cell, uint32_t day, double &z0, double &z1) {
  uint32_t c[4] = {cell, day, 0x424f584du, 0u};
  philox(c, uint32_t(seed), uint32_t(seed >> 32));
  const double u1 = (double(c[0]) + 0.5) * 2.3283064365386963e-10;
  const double u2 = (double(c[1]) + 0.5) * 2.3283064365386963e-10;
  const double r = sqrt(-2.0 * wg_log(u1));
  double sn, cs;
  wg_sincos2pi(u2, &sn, &cs);
  z0 = r * cs;
  z1 = r * sn;
}

// Device accessors of the routing links and chain constants ([k][N] x 16 bytes).
__device__ inline WgChainLink load_link(const WgStatic &S, int p) {
  const WgI4 a = S.clink[p], b = S.clink[int64_t(S.n) + p];
  WgChainLink lk;
  lk.flags = a.x;
  lk.station = a.y;
  lk.nu = int16_t(a.z & 0xffff);
  lk.nu_level = int16_t(a.z >> 16);
  lk.level = a.w;
  const int32_t u[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    lk.up[2 * k] = int16_t(uint32_t(u[k]) & 0xffff);
    lk.up[2 * k + 1] = int16_t(uint32_t(u[k]) >> 16);
  }
  return lk;
}

__device__ inline WgChainStatic load_cstat(const WgStatic &S, int p) {
  const int64_t n = S.n;
  const WgD2 a = S.cstat[p], b = S.cstat[n + p], c = S.cstat[2 * n + p];
  return WgChainStatic{a.x, a.y, b.x, b.y, c.x, c.y};
}

// Pair 0 of the routing record of (member m, cell p); pair k is at [k * N].
__device__ inline WgD2 *rec_at(const WgState &X, int n, int m, int p) {
  return X.crec + int64_t(m) * (WG_DEV_NREC / 2) * n + p;
}

// Writes what chain_step changed.
__device__ inline void store_chain(const WgState &X, int64_t mp, int flags, const ChainOut &o) {
  if (flags & CFL_LAKE) X.glo_lake[mp] = o.glo_lake;
  if (flags & CFL_RES) {
    X.res[mp] = o.res;
    X.k_release[mp] = o.k_release;
  }
  if (flags & CFL_GWET) X.glo_wetl[mp] = o.glo_wetl;
  if (flags & CFL_ARID) X.gw[mp] = o.gw;
  if ((flags & CFL_ARID) && X.pot_lgw) X.pot_lgw[mp] = o.lgw;
  X.river[mp] = o.river;
  X.tv[mp] = o.tv;
  if (X.use_rem) X.use_rem[mp] = o.rem;
}

// updateNetAbstractionGW, routing.cpp:5503-5572: the day's net abstraction from
// groundwater, with the return flows of irrigation from surface water reduced
// when the last day's surface-water use went unsatisfied (G_dailyRemainingUse > 0)
// or given back when a delayed use was satisfied (< 0).  U points at the cell's
// first water-use field, us is the stride between fields.
__device__ inline double update_net_abstraction_gw(double *U, int64_t us, double frgi, double nug) {
  double &dru = U[WG_USE_DAILY_REMAINING * us], &ui = U[WG_USE_UNSAT_IRRIG * us];
  double &uo = U[WG_USE_UNSAT_OTHER * us], &rrf = U[WG_USE_REDUCED_RETURN * us];
  const double wu = U[WG_USE_WITHDRAWAL_IRRIG * us], cu = U[WG_USE_CONSUMPTIVE_IRRIG * us];
  if (dru > 1.e-12) {
    if (wu > 0.) {
      const double eff = cu / wu;
      const double factor = 1 - (1 - frgi) * (1 - eff);
      double WUsiNew = 1 / factor * (wu * factor - dru);
      if (WUsiNew < 0.) {
        WUsiNew = 0.;
        ui += wu * factor;
        uo += dru - (wu * factor);
      } else {
        ui += dru;
      }
      const double returnflowChange = (frgi * (1 - eff) * (WUsiNew - wu));
      rrf += returnflowChange;
      dru = 0.;
      return nug - returnflowChange;
    }
    uo += dru;
    return nug;
  }
  if (dru < -1.e-12) {
    double fromIrrig = dru + uo;
    if (fromIrrig < 0.) {
      uo = 0.;
    } else {
      uo += dru;
      dru = 0.;
      return nug;
    }
    if (ui == 0.) {
      dru = 0.;
      return nug;
    }
    double ratio = (fromIrrig / ui);
    if (ratio < -1.) {
      ratio = -1.;
      fromIrrig = ui * -1.;
    }
    const double returnflowChange = (ratio * rrf);
    ui += fromIrrig;
    rrf += returnflowChange;
    dru = 0.;
    return nug - returnflowChange;
  }
  return nug;
}

// annualWaterUsePostProcessing, routing.cpp:5257-5296, for one cell
__device__ inline void water_use_year_end(double *U, int64_t us) {
  double &tu = U[WG_USE_TOTAL_UNSAT * us], &ui = U[WG_USE_UNSAT_IRRIG * us];
  double &uo = U[WG_USE_UNSAT_OTHER * us], &rrf = U[WG_USE_REDUCED_RETURN * us];
  double &tu_py = U[WG_USE_UNSAT_PREV_YEAR * us], &ui_py = U[WG_USE_UNSAT_IRRIG_PREV_YEAR * us];
  double &uo_py = U[WG_USE_UNSAT_OTHER_PREV_YEAR * us], &rrf_py = U[WG_USE_REDUCED_RETURN_PREV_YEAR * us];
  if (tu < tu_py) {
    tu_py = tu;
    ui_py = ui;
    uo_py = uo;
    rrf_py = rrf;
    ui = 0.;
    uo = 0.;
    rrf = 0.;
    tu = 0.;
  } else {
    tu -= tu_py;
    ui -= ui_py;
    uo -= uo_py;
    rrf -= rrf_py;
    if ((ui + uo) < 0) {
      ui = 0;
      uo = 0;
      rrf = 0;
    } else if (uo < 0) {
      ui += uo;
      uo = 0;
    } else if (ui < 0) {
      uo += ui;
      ui = 0;
      rrf = 0;
    }
    if (rrf > 0) {
      rrf = 0;
      ui = 0;
    }
  }
}

// routing.cpp:7309-7341 getNewRiverVelocity
// c1 = 1. / (mult * rough) and c2 = pow(slope, 0.5) are per-cell constants computed
// on the host with the same operations, so the product below is the reference's.
__device__ inline double new_river_velocity(double c1, double c2, double bottom, double stor, double length) {
  const double csa = (stor / length) * 1000000.;
  const double depth = -bottom / 4.0 + sqrt(bottom * bottom / 16.0 + 0.5 * csa);
  const double wp = bottom + 2.0 * depth * sqrt(5.0);
  const double hr = csa / wp;
  double v = c1 * wg_pow(hr, (2. / 3.)) * c2;
  v = v * 86.4;
  return (v < 0.00001) ? 0.00001 : v;
}

// LAI state machine, lai.cpp:179-293
__device__ inline double lai_day(int &days, int init_days, int &status, int lct, bool arid, double lmin,
                                 double lmax, double &precsum, double prec, bool growing) {
  if (growing) {
    if (status == 0) {
      if (days >= init_days) {
        days++;
        precsum += prec;
        if (precsum > 40.) {
          if (days >= init_days + 30) {
            days = init_days + 30;
            status = 1;
          }
          return (lmin + (lmax - lmin) * (days - init_days) / 30.);
        }
        days = init_days;
        return lmin;
      }
      days++;
      precsum += prec;
      return lmin;
    }
    if (days <= 30) {
      days--;
      if (lct <= 2) status = 0;
      if (days <= 0) {
        days = 0;
        status = 0;
        precsum = 0.;
      }
      return (lmax - (lmax - lmin) * (30 - days) / 30.);
    }
    if (arid && (prec < 0.5))
      days--;
    else
      days = 30 + init_days;
    return lmax;
  }
  if (status == 0) {
    if (days > init_days) {
      days++;
      precsum += prec;
      if (precsum > 40.) {
        if (days >= init_days + 30) {
          days = init_days + 30;
          status = 1;
        }
        return (lmin + (lmax - lmin) * (days - init_days) / 30.);
      }
      days = init_days;
      return lmin;
    }
    precsum += prec;
    return lmin;
  }
  if (days <= 30) {
    days--;
    if (days <= 0) {
      days = 0;
      status = 0;
      precsum = 0.;
    }
    return (lmax - (lmax - lmin) * (30 - days) / 30.);
  }
  days--;
  return lmax;
}

// Finish the previous day for one cell: reduction factors of the global
// storages, river width and area fraction (routing.cpp:2792-2802, 3071-3078,
// 3286-3295, 3546-3586), then the surface-water and land area fractions
// (routing.cpp:5046-5174) and updateLandAreaFrac (routing.cpp:5343-5352).
struct PostDay {
  double laf, laf_prev, red_ll, red_lw, red_gw, red_gl, red_res, river_frac_next;
};

// With water use, the finished day's end of the routing loop comes first
// (routing.cpp:3590-3622, 3898-3919): the use the river left is taken from the
// local lake, whose area reduction factor is recomputed, the unsatisfied use is
// kept and the month's irrigation withdrawal and consumptive use are recorded
// (fin_ym: model year x 12 + month of the finished day); after day 365 the
// year-end processing follows (fin_year_end).
template <bool kUse>
__device__ void post_day(const WgStatic &S, const WgState &X, int64_t mp, int p, PostDay &out, int fin_ym,
                         bool fin_year_end) {
  // Every input is loaded before the first store, so that the loads are in flight
  // together (the arrays cover all cells, so the loads of unused fields are in bounds).
  const double evarexp = S.evarexp[p];
  const double cellArea = S.cell_area[p];
  const double lake_area = S.lake_area[p], lake_depth = S.lake_depth[p], glo_lake = X.glo_lake[mp];
  const double reservoir_area = S.reservoir_area[p], stor_cap = S.stor_cap[p], res = X.res[mp];
  const double glo_wetland = S.glo_wetland[p], wetl_depth = S.wetl_depth[p], glo_wetl = X.glo_wetl[mp];
  const bool contcell = S.contcell[p] != 0;
  const double L = S.river_length[p];
  const double rw_c0 = S.rw_c0[p], rw_c1 = S.rw_c1[p], rw_c2 = S.rw_c2[p], rw_c3 = S.rw_c3[p];
  const double st = X.river[mp], smax = S.river_stor_max[p];
  const double loc_lake = S.loc_lake[p], loc_wetland = S.loc_wetland[p];
  const double f_glo_lake = S.f_glo_lake[p], contfreq = S.contfreq[p];
  double red_ll = X.red_ll[mp], red_lw = X.red_lw[mp], red_gw = X.red_gw[mp];
  double river_frac_next = X.river_frac_next[mp];
  double river_frac_change = X.river_frac_change[mp];
  double red_river = X.red_river[mp];
  const double fswb_old = X.fswb_next[mp];
  const double laf = X.laf[mp];
  if (kUse && contcell) {
    const int64_t us = S.use_stride;
    double *U = X.use + mp;
    double rem = X.use_rem[mp];
    if ((loc_lake > 0.) && (rem > 0.)) {
      const double maxStorage = ((loc_lake) / 100.) * cellArea * lake_depth;
      double ls = X.loc_lake[mp];
      if (ls > (-1.) * maxStorage) {
        if (rem < (maxStorage + ls)) {
          ls -= rem;
          rem = 0.;
        } else {
          rem -= (maxStorage + ls);
          ls = (-1.) * maxStorage;
        }
      }
      double r = 1. - wg_pow(fabs(ls - maxStorage) / (2. * maxStorage), (evarexp * kEvapoRedExp));
      if (r < 0.) r = 0.;
      if (r > 1.) r = 1.;
      red_ll = r;
      X.loc_lake[mp] = ls;
    }
    if (S.use_delayed == 1) {
      const double tu = U[WG_USE_TOTAL_UNSAT * us];
      U[WG_USE_PREV_TOTAL_UNSAT * us] = tu;
      U[WG_USE_DAILY_REMAINING * us] = rem - tu;
      U[WG_USE_TOTAL_UNSAT * us] = rem;
    } else {
      U[WG_USE_TOTAL_UNSAT * us] += rem;
      U[WG_USE_DAILY_REMAINING * us] = rem;
    }
    const int64_t t = int64_t(fin_ym) * S.n + p;
    U[WG_USE_WITHDRAWAL_IRRIG * us] = S.use_wu[t];
    U[WG_USE_CONSUMPTIVE_IRRIG * us] = S.use_cu[t];
    X.use_rem[mp] = 0.;
    if (fin_year_end) water_use_year_end(U, us);
  }

  double red_gl = 0., red_res = 0.;
  if (lake_area > 0.) {
    const double maxStorage = lake_area * lake_depth;
    double r = 1. - wg_pow(fabs(glo_lake - maxStorage) / (2. * maxStorage), (evarexp * kEvapoRedExp));
    if (r < 0.) r = 0.;
    if (r > 1.) r = 1.;
    red_gl = r;
  }
  if (reservoir_area > 0.) {
    const double maxStorage = stor_cap;
    double r = 1. - wg_pow(fabs(res - maxStorage) / maxStorage, kEvapoRedExpRes);
    if (r < 0.) r = 0.;
    if (r > 1.) r = 1.;
    red_res = r;
  }
  if (glo_wetland > 0) {
    const double maxStorage = ((glo_wetland) / 100.) * cellArea * wetl_depth;
    double r = 1. - wg_pow(fabs(glo_wetl - maxStorage) / maxStorage, (evarexp * kEvapoRedExp));
    if (r < 0.) r = 0.;
    if (r > 1.) r = 1.;
    red_gw = r;
  }
  if (contcell) {
    const double riverAreaFrac = river_frac_next;
    const double csa = st / L;
    // -B / (4. * 1000.), B / 1000. * B / (16. * 1000.), B / 1000. and width_bf / 1000.
    // are per-cell constants computed on the host with these operations
    const double depth = rw_c0 + sqrt(rw_c1 + 0.5 * csa);
    double width = rw_c2 + 4. * depth;
    if (width > rw_c3) width = rw_c3;
    red_river = 1. - wg_pow(fabs(st - smax) / smax, (evarexp * kEvapoRedExp));
    if (red_river < 0.) red_river = 0.;
    if (red_river > 1.) red_river = 1.;
    river_frac_next = red_river * L * width * 100. / cellArea;
    river_frac_change = river_frac_next - riverAreaFrac;
  }
  // fractions for the next day
  double fLocLake, fLocWet, fGloWet;
  if ((loc_lake > 0.) && (red_ll > 0.))
    fLocLake = (red_ll * loc_lake / 100.);
  else
    fLocLake = 0.;
  if ((loc_wetland > 0.) && (red_lw > 0.))
    fLocWet = (red_lw * loc_wetland / 100.);
  else
    fLocWet = 0.;
  if ((glo_wetland > 0.) && (red_gw > 0.))
    fGloWet = (red_gw * glo_wetland / 100.);
  else
    fGloWet = 0.;
  double fswb_next = fLocLake + fLocWet + fGloWet;
  const double max_raf = contfreq / 100. - f_glo_lake;
  if (((lake_area > 0.) || (reservoir_area > 0.)) && (f_glo_lake == 1.)) {
    river_frac_next = 0.;
    river_frac_change = 0.;
    red_river = 0.;
  } else {
    if (river_frac_next <= max_raf) {
      if (fswb_next > (max_raf - river_frac_next)) {
        const double corr = (max_raf - river_frac_next) / fswb_next;
        red_ll *= corr;
        red_lw *= corr;
        red_gw *= corr;
        if ((fLocLake > 0.) && (red_ll > 0.)) {
          fLocLake = (red_ll * loc_lake / 100.);
        } else {
          red_ll = 0.;
          fLocLake = 0.;
        }
        if ((fLocWet > 0.) && (red_lw > 0.)) {
          fLocWet = (red_lw * loc_wetland / 100.);
        } else {
          red_lw = 0.;
          fLocWet = 0.;
        }
        if ((fGloWet > 0.) && (red_gw > 0.)) {
          fGloWet = (red_gw * glo_wetland / 100.);
        } else {
          red_gw = 0.;
          fGloWet = 0.;
        }
      }
    } else {
      const double deficit = river_frac_next - max_raf;
      river_frac_change -= deficit;
      red_river *= max_raf / river_frac_next;
      river_frac_next = max_raf;
      fLocLake = 0.;
      fLocWet = 0.;
      fGloWet = 0.;
    }
  }
  fswb_next = fLocLake + fLocWet + fGloWet;
  const double change_pct = fswb_next * 100. - fswb_old * 100.;
  double laf_next = laf - (change_pct + (river_frac_change * 100.));
  if (laf_next < 0.) laf_next = 0.;
  if (lake_area > 0.) X.red_gl[mp] = red_gl;
  if (reservoir_area > 0.) X.red_res[mp] = red_res;
  X.laf_prev[mp] = laf;
  X.laf[mp] = laf_next;
  X.fswb_next[mp] = fswb_next;
  X.red_ll[mp] = red_ll;
  X.red_lw[mp] = red_lw;
  X.red_gw[mp] = red_gw;
  X.red_river[mp] = red_river;
  X.river_frac_next[mp] = river_frac_next;
  X.river_frac_change[mp] = river_frac_change;
  out.laf = laf_next;
  out.laf_prev = laf;
  out.red_ll = red_ll;
  out.red_lw = red_lw;
  out.red_gw = red_gw;
  out.red_gl = red_gl;
  out.red_res = red_res;
  out.river_frac_next = river_frac_next;
}


// ---- one elevation sub-cell of the snow loop, daily.cpp:920-1044 ----

struct SnowDay {
  double tempC, gradnt, scale_num, laf, prec_to_soil, soil_pet, snowfz, snowmt, melt_factor;
  double eff_bare;  // the effective precipitation of a sub-cell without snow above freezing
  double inv_laf;   // 1 / laf, correctly rounded
  int elev0;
  bool melt_ok;     // melt_factor >= 0 (a NaN or negative factor takes the full step)
};

struct SnowAcc {
  double snow_sum, eff, evapo, change, temp_max;
  int thresh;
};

// Updates one sub-cell storage s and the cell's running sums in sub-cell order
// (daily.cpp:920-1044); temp_elev is the sub-cell temperature T - (z_e - z0) gradnt
// and melt_factor = M_DEGDAY_F * ddf_lct (daily.cpp:1013).  The reference's branches
// are selects here: every operation is the reference's, on the same operands, and
// a branch not taken adds +0 to a sum, which leaves it unchanged (the sums start at
// +0 and never hold -0).  Without branches the two sub-cells of a pair are
// independent chains that the compiler interleaves.
__device__ inline double snow_subcell(double s, double temp_elev, int elev_e, bool first, const SnowDay &d,
                                      SnowAcc &a) {
  {
    // s * scale_num / laf with the reused reciprocal: q0 = x (1/laf), one exact
    // residual and one correction give the correctly rounded quotient (Markstein's
    // theorem for a correctly rounded reciprocal; tests/test_division.py compares
    // it with the division on random operands).  An infinite or NaN x is its own quotient by the
    // finite positive laf.  (The quotient of a finite x overflows only for
    // x > 1e292 mm, since a positive land area fraction is at least about 1e-16.)
    const double x = s * d.scale_num;
    const double q0 = x * d.inv_laf;
    const double q = fma(fma(-q0, d.laf, x), d.inv_laf, q0);
    s = (fabs(x) <= 1.7976931348623157e308) ? q : x;
  }
  s = (fabs(s) <= kMinStorVol) ? 0. : s;
  const double init_s = s;
  // threshold elevation above 1000 mm of storage (daily.cpp:958-976)
  const bool big = s > 1000.;
  const bool use_thresh = big && a.thresh > 0;
  const double temp_thresh = d.tempC - ((a.thresh - d.elev0) * d.gradnt);
  a.thresh = (big && a.thresh == 0) ? elev_e : a.thresh;
  temp_elev = use_thresh ? temp_thresh : temp_elev;
  // snowfall and sublimation below the freezing temperature, rain above it
  const bool freeze = temp_elev <= d.snowfz;
  const double s_fall = s + d.prec_to_soil;
  const bool keeps = s_fall > d.soil_pet;
  const double s_subl = s_fall - d.soil_pet;
  a.evapo += freeze ? (keeps ? d.soil_pet : s_fall) : 0.;
  s = freeze ? (keeps ? s_subl : 0.) : s;
  const double eff_before = freeze ? 0. : d.prec_to_soil;
  // degree-day melt
  const bool melts = (temp_elev > d.snowmt) && !(s < 0.);
  const double melt_pot = d.melt_factor * (temp_elev - d.snowmt);
  const bool all = melt_pot > s;
  const double s_left = s - melt_pot;
  const double snowmelt_elev = melts ? (all ? s : melt_pot) : 0.;
  s = melts ? (all ? 0. : s_left) : s;
  const double eff_elev = eff_before + snowmelt_elev;
  a.change += s - init_s;
  if (first) a.temp_max = temp_elev;
  a.snow_sum += s;
  a.eff += eff_elev;
  return s;
}

}  // namespace

// Per-cell inputs of the cell kernel: statics (s_), state (x_) and land-cover
// table entries (t_).  The kernel loads them at two points, each before the first
// store of the part that reads them, so that a part's loads are in flight together
// (after a store the compiler must keep later loads behind it): load_cell_in before
// post_day for PET, LAI, canopy and snow, load_cell_in_b after the snow loop for
// soil and routing.  Every array covers all cells, so the loads of fields a cell
// does not use are in bounds.
struct CellIn {
  int16_t s_aindex;
  double s_alpha;
  double s_bottom_width;
  float s_built_up;
  double s_cell_area;
  double s_contfreq;
  double s_degday;
  double s_e_gw;
  double s_e_sw;
  int16_t s_elev0;
  double s_evarexp;
  double s_fswb_init;
  double s_glo_wetland;
  float s_gw_factor;
  double s_inv_kgw;
  double s_k_sw;
  float s_lai_max;
  double s_lake_area;
  double s_lake_depth;
  int8_t s_land_cover;
  int8_t s_ldd;
  double s_loc_lake;
  double s_loc_wetland;
  double s_m_netrad;
  double s_m_prec;
  double s_mcwh;
  double s_mean_outflow, s_prov_rel;
  double s_ome_gw;
  double s_pcritgwa;
  double s_pet_mxdy;
  double s_reservoir_area;
  int16_t s_rgmax;
  double s_river_length;
  double s_vel_c1;
  double s_vel_c2;
  double s_stor_cap;
  double s_t_gradnt;
  double s_t_snowfz;
  double s_t_snowmt;
  int8_t s_texture;
  double s_wetl_depth;
  double x_canopy;
  double x_cfa;
  double x_gamma;
  double x_glo_lake;
  double x_glo_wetl;
  double x_gw;
  double x_k_release;
  int32_t x_lai_days;
  double x_lai_precsum;
  int32_t x_lai_status;
  double x_loc_lake;
  double x_loc_wetl;
  double x_res;
  double x_river;
  float x_smax;
  double x_snow;
  uint8_t x_snow_any;
  double x_soil;
  double x_transfer;
  double t_albedo, t_albedo_snow, t_ddf, t_emissivity;
  float t_lai_a, t_lai_b;
  int32_t t_initial_days;
  WgChainLink link;
  WgChainStatic cstat;
};

__device__ inline void load_cell_in(const WgStatic &S, const WgState &X, const WgTables &T, int64_t mp, int p,
                                    CellIn &ci) {
  ci.s_aindex = S.aindex[p];
  ci.s_alpha = S.alpha[p];
  ci.s_degday = S.degday[p];
  ci.s_elev0 = S.elev0[p];
  ci.s_lai_max = S.lai_max[p];
  ci.s_land_cover = S.land_cover[p];
  ci.s_m_netrad = S.m_netrad[p];
  ci.s_m_prec = S.m_prec[p];
  ci.s_mcwh = S.mcwh[p];
  ci.s_pet_mxdy = S.pet_mxdy[p];
  ci.s_t_gradnt = S.t_gradnt[p];
  ci.s_t_snowfz = S.t_snowfz[p];
  ci.s_t_snowmt = S.t_snowmt[p];
  ci.x_canopy = X.canopy[mp];
  ci.x_lai_days = X.lai_days[mp];
  ci.x_lai_precsum = X.lai_precsum[mp];
  ci.x_lai_status = X.lai_status[mp];
  ci.x_snow = X.snow[mp];
  ci.x_snow_any = X.snow_any[mp];
  ci.x_transfer = X.transfer[mp];
  const int l = ci.s_land_cover - 1;
  ci.t_albedo = T.albedo[l];
  ci.t_albedo_snow = T.albedo_snow[l];
  ci.t_ddf = T.ddf[l];
  ci.t_emissivity = T.emissivity[l];
  ci.t_initial_days = T.initial_days[l];
  ci.t_lai_a = T.lai_a[l];
  ci.t_lai_b = T.lai_b[l];
}

__device__ inline void load_cell_in_b(const WgStatic &S, const WgState &X, int64_t mp, int p, CellIn &ci) {
  ci.s_bottom_width = S.bottom_width[p];
  ci.s_built_up = S.built_up[p];
  ci.s_cell_area = S.cell_area[p];
  ci.s_contfreq = S.contfreq[p];
  ci.s_e_gw = S.e_gw[p];
  ci.s_e_sw = S.e_sw[p];
  ci.s_evarexp = S.evarexp[p];
  ci.s_fswb_init = S.fswb_init[p];
  ci.s_glo_wetland = S.glo_wetland[p];
  ci.s_gw_factor = S.gw_factor[p];
  ci.s_inv_kgw = S.inv_kgw[p];
  ci.s_k_sw = S.k_sw[p];
  ci.s_lake_area = S.lake_area[p];
  ci.s_lake_depth = S.lake_depth[p];
  ci.s_ldd = S.ldd[p];
  ci.s_loc_lake = S.loc_lake[p];
  ci.s_loc_wetland = S.loc_wetland[p];
  ci.s_mean_outflow = S.mean_outflow[p];
  ci.s_prov_rel = S.res_prov_rel[p];
  ci.s_ome_gw = S.ome_gw[p];
  ci.s_pcritgwa = S.pcritgwa[p];
  ci.s_reservoir_area = S.reservoir_area[p];
  ci.s_rgmax = S.rgmax[p];
  ci.s_river_length = S.river_length[p];
  ci.s_vel_c1 = S.vel_c1[p];
  ci.s_vel_c2 = S.vel_c2[p];
  ci.s_stor_cap = S.stor_cap[p];
  ci.s_texture = S.texture[p];
  ci.s_wetl_depth = S.wetl_depth[p];
  ci.x_cfa = X.cfa[mp];
  ci.x_gamma = X.gamma[mp];
  ci.x_glo_lake = X.glo_lake[mp];
  ci.x_glo_wetl = X.glo_wetl[mp];
  ci.x_gw = X.gw[mp];
  ci.x_k_release = X.k_release[mp];
  ci.x_loc_lake = X.loc_lake[mp];
  ci.x_loc_wetl = X.loc_wetl[mp];
  ci.x_res = X.res[mp];
  ci.x_river = X.river[mp];
  ci.x_smax = X.smax[mp];
  ci.x_soil = X.soil[mp];
  ci.link = load_link(S, p);
  ci.cstat = load_cstat(S, p);
}

struct WgClock {
  int64_t day0;   // forcing day (index into the loaded forcing) of launch offset 0
  int32_t doy0;   // julian day of launch offset 0
  int32_t started;
};

// One lane per (member, cell).  Grid: x over cell tiles, y over members.
// kSnowChunk: sub-cell pairs per double-buffered chunk of the snow loop (5 or 10).
// With chunks of 5 pairs the kernel is held to 168 registers, three wavefronts per
// SIMD, for many members, where it is bandwidth-bound; chunks of 10 leave the
// compiler free, for one member, where it is latency-bound.
// kUse: water use on (the handle's tables are set); the instantiation without it
// is the model without water use, with no use code at all.
template <int kSnowChunk, bool kUse>
__global__ __launch_bounds__(256, kSnowChunk == 5 ? 3 : 1) void wg_cell_kernel(WgStatic S, WgState X, WgTables T,
                                                      const float *__restrict__ forcing, int64_t n_forcing_days,
                                                      const WgClock *__restrict__ clock, int k,
                                                      int do_post, int do_vertical, int write_diag,
                                                      int daily_slot) {
  load_exp2_table();
  const int t0 = S.tile_start[blockIdx.x];
  if (t0 < 0) return;
  const int p = t0 + int(threadIdx.x);
  if (p >= S.tile_end[blockIdx.x]) return;
  const int m = blockIdx.y;
  const int n = S.n;
  const int64_t mp = int64_t(m) * n + p;

  CellIn ci;
  if (do_vertical) load_cell_in(S, X, T, mp, p, ci);
  // the fractions post_day writes, which the rest of the kernel reads
  PostDay pd;
  if (do_post) {
    // the finished day is day k - 1 of the block (the closing pass runs with k = days)
    int fin_ym = 0;
    bool fin_year_end = false;
    if (kUse) {
      const int64_t fin = clock->day0 + k - 1;
      const int fin_doy = int((clock->doy0 - 1 + k - 1) % 365) + 1;
      fin_ym = int((fin - S.use_day0) / 365) * 12 + day_info(fin_doy).month;
      fin_year_end = fin_doy == 365;
    }
    post_day<kUse>(S, X, mp, p, pd, fin_ym, fin_year_end);
  } else if (do_vertical) {
    pd.laf = X.laf[mp];
    pd.laf_prev = X.laf_prev[mp];
    pd.red_ll = X.red_ll[mp];
    pd.red_lw = X.red_lw[mp];
    pd.red_gw = X.red_gw[mp];
    pd.red_gl = X.red_gl[mp];
    pd.red_res = X.red_res[mp];
    pd.river_frac_next = X.river_frac_next[mp];
  }
  // calcNewDay and routing run for continental cells only (integrateWGHM.cpp:770,
  // routing.cpp:1876); other cells keep a zero transported volume.
  if (!do_vertical) return;
  if (!S.contcell[p]) {
    const int st = S.station[p];
    if (st >= 0) X.st_daily[(int64_t(m) * WG_DAILY_SLOTS + daily_slot) * S.n_stations + st] = 0.;
    return;
  }

  const int64_t fday = clock->day0 + k;
  const int doy = int((clock->doy0 - 1 + k) % 365) + 1;
  const bool started = (clock->started != 0) || (k > 0) || (do_post != 0);
  if (fday < 0 || fday >= n_forcing_days) return;
  const float *f = forcing + fday * 4 * int64_t(n);
  float tf = f[p], pf = f[n + p];
  const float swf = f[2 * n + p], lwf = f[3 * n + p];
  if (X.perturb) {
    const uint64_t seed = X.member_seed[m];
    const uint32_t ref = uint32_t(S.ref_index[p]);
    double zp, zt;
    normal_pair(seed, ref, uint32_t(fday), zp, zt);
    pf = float(double(pf) * wg_exp(X.prec_sigma * zp));
    tf = float(double(tf) + X.temp_sigma * zt);
  }
  // water use of the day: the desired surface-water use (routing.cpp:2193-2290)
  // and the net abstraction from groundwater (routing.cpp:7440-7441, adapted by
  // updateNetAbstractionGW in every groundwater branch when the last day left a
  // remaining use; the adaptation reads only the last day's state, so it is done
  // here for (semi-)arid cells too, whose groundwater step the chain runs)
  double use_desired = 0., use_nag = 0.;
  int64_t use_t = 0;
  if (kUse) {
    use_t = (int64_t((fday - S.use_day0) / 365) * 12 + day_info(doy).month) * n + p;
    const int64_t us = S.use_stride;
    double *U = X.use + mp;
    use_nag = S.use_nug[use_t];
    if (U[WG_USE_DAILY_REMAINING * us] != 0.) use_nag = update_net_abstraction_gw(U, us, S.use_frgi[p], use_nag);
    use_desired = S.use_nus[use_t];
    if (S.use_delayed == 1) use_desired += U[WG_USE_TOTAL_UNSAT * us];
  }

  // ------------------------------------------------------------------
  // calcNewDay, daily.cpp:94-1279
  // ------------------------------------------------------------------
  const int lct = ci.s_land_cover;
  const bool arid_gw = (ci.s_aindex == 1);
  const double landAreaFrac = pd.laf;
  const double lafPrev = started ? pd.laf_prev : landAreaFrac;  // daily.cpp:159-169
  double dailyTempC = tf, dailyPrec = pf;
  const double dailyShortWave = swf, dailyLongWave = lwf;
  dailyPrec = ci.s_m_prec * dailyPrec;  // daily.cpp:248
  const double temp2 = dailyTempC + 237.3;
  const double e_s = 0.6108 * wg_exp(17.27 * dailyTempC / temp2);
  const double alpha = ci.s_alpha;
  const double maxDailyPET = ci.s_pet_mxdy;

  // lai.cpp:152-177
  double dailyLai;
  {
    const float lmax_f = ci.s_lai_max;
    const double lmax = double(lmax_f);
    const double lmin = ci.t_lai_a + ci.t_lai_b * lmax_f;
    int days = ci.x_lai_days, status = ci.x_lai_status;
    double precsum = ci.x_lai_precsum;
    dailyLai = lai_day(days, ci.t_initial_days, status, lct, arid_gw, lmin, lmax, precsum, dailyPrec,
                       dailyTempC > 8.);
    X.lai_days[mp] = days;
    X.lai_status[mp] = status;
    X.lai_precsum[mp] = precsum;
  }
  // daily.cpp:366-462 albedo, latent heat, radiation (cloud = 1)
  const double albedo = (ci.x_snow > 3.) ? ci.t_albedo_snow : ci.t_albedo;
  const double lat_heat = (dailyTempC > 0) ? (2.501 - 0.002361 * dailyTempC) : 2.835;
  const double conv_Wm2_to_mmd = 0.0864 / lat_heat;
  const double solar_rad = conv_Wm2_to_mmd * dailyShortWave;
  const double long_wave_rad_in = conv_Wm2_to_mmd * dailyLongWave;
  const double temp_K = dailyTempC + 273.2;
  const double long_wave_rad_out = ci.t_emissivity * 0.000000004903 * wg_pow(temp_K, 4.) / lat_heat;
  const double net_long_wave_rad = long_wave_rad_in - long_wave_rad_out;
  const double net_short_wave_rad = solar_rad * (1. - albedo);
  const double net_rad = ci.s_m_netrad * (net_short_wave_rad + net_long_wave_rad);
  const double openWaterNetRad = solar_rad * (1. - kOpenWaterAlbedo) + net_long_wave_rad;
  // daily.cpp:476-522 Priestley-Taylor
  const double inc_svp = 4098. * e_s / (temp2 * temp2);
  const double gamma = (0.0016286 * 101.3) / lat_heat;
  const double dailyPET = (net_rad <= 0.) ? 0. : alpha * (inc_svp * net_rad) / (inc_svp + gamma);
  const double dailyOpenWaterPET =
      (openWaterNetRad <= 0.) ? 0. : alpha * (inc_svp * openWaterNetRad) / (inc_svp + gamma);

  double landStorageChangeSum = 0., initialStorage = 0.;
  double dailyCanopyEvapo = 0., daily_prec_to_soil = 0., dailySoilPET = 0.;
  double dailySnowEvapo = 0., dailyEffPrec = 0., immediate_runoff = 0., dailyAET = 0.;
  double daily_runoff = 0., total_daily_runoff = 0., daily_gw_recharge = 0., pot_gw_recharge = 0.;
  double soil_water_overflow = 0., neg_land_aet = 0.;
  double dailyCellLandAET_uncorr = 0.;
  double transfer = ci.x_transfer;

  // interception, daily.cpp:825-895
  double canopy = ci.x_canopy;
  if (landAreaFrac <= 0.) {
    transfer = canopy;
    canopy = 0.;
  } else {
    canopy *= lafPrev / landAreaFrac;
    if (fabs(canopy) <= kMinStorVol) canopy = 0.;
    initialStorage = canopy;
    if (dailyLai > 0.00001) {
      const double max_canopy_storage = ci.s_mcwh * dailyLai;
      const double canopy_deficiency = max_canopy_storage - canopy;
      if (dailyPrec < canopy_deficiency) {
        canopy += dailyPrec;
        daily_prec_to_soil = 0.;
      } else {
        canopy = max_canopy_storage;
        daily_prec_to_soil = dailyPrec - canopy_deficiency;
      }
      const double canopy_water_content = canopy;
      dailyCanopyEvapo = dailyPET * wg_pow((canopy_water_content / max_canopy_storage), kCanopyEvapoExp);
      if (dailyCanopyEvapo > canopy_water_content) {
        dailyCanopyEvapo = canopy_water_content;
        dailySoilPET = dailyPET - canopy_water_content;
        canopy = 0.0;
      } else {
        canopy -= dailyCanopyEvapo;
        dailySoilPET = dailyPET - dailyCanopyEvapo;
      }
    } else {
      daily_prec_to_soil = dailyPrec;
      dailySoilPET = dailyPET;
      dailyCanopyEvapo = 0.0;
    }
    landStorageChangeSum += canopy - initialStorage;
  }
  X.canopy[mp] = canopy;
  if (dailySoilPET < 0.) dailySoilPET = 0.0;

  // snow on the elevation sub-cells, daily.cpp:897-1062, serial in sub-cell order.
  // Sub-cells are stored in pairs ([member][pair][cell] x 2 doubles, elevations as
  // packed int16 pairs) and streamed in double-buffered chunks so that a
  // wavefront keeps many 16-byte loads in flight.
  const double T_SNOWFZ = ci.s_t_snowfz, T_SNOWMT = ci.s_t_snowmt;
  const double T_GRADNT = ci.s_t_gradnt, DEGDAY_F = ci.s_degday;
  const double ddf = ci.t_ddf;
  SnowAcc acc;
  acc.snow_sum = 0.;
  acc.eff = 0.;
  acc.evapo = 0.;
  acc.change = 0.;
  acc.temp_max = 0.;
  acc.thresh = 0;
  const int elev0 = ci.s_elev0;
  WgD2 *sp = X.snow2 + int64_t(m) * WG_DEV_NPAIR * n + p;
  const int32_t *ep = S.elev_pairs + p;
  // A cell whose sub-cell storages are all exactly zero (snow_any == 0) starts the
  // loop from zeros without reading them, and stores only storages that become
  // non-zero: on warm days such a cell moves no snow bytes at all.
  const bool snow_any = ci.x_snow_any != 0;
  if (landAreaFrac <= 0.) {
    for (int k = 0; k < WG_DEV_NPAIR; k++) {
      const WgD2 v = snow_any ? sp[int64_t(k) * n] : WgD2{0., 0.};
      transfer += v.x / 100.;
      transfer += v.y / 100.;
      if (snow_any) sp[int64_t(k) * n] = WgD2{0., 0.};
    }
    X.snow_any[mp] = 0;
  } else {
    SnowDay sd;
    sd.tempC = dailyTempC;
    sd.elev0 = elev0;
    sd.gradnt = T_GRADNT;
    sd.scale_num = lafPrev;
    sd.laf = landAreaFrac;
    sd.prec_to_soil = daily_prec_to_soil;
    sd.soil_pet = dailySoilPET;
    sd.snowfz = T_SNOWFZ;
    sd.snowmt = T_SNOWMT;
    sd.melt_factor = DEGDAY_F * ddf;
    sd.eff_bare = daily_prec_to_soil + 0.;
    sd.inv_laf = 1. / landAreaFrac;
    sd.melt_ok = sd.melt_factor >= 0.;
    bool any_after = false;
    WgD2 sa[kSnowChunk], sb[kSnowChunk];
    int32_t ea[kSnowChunk], eb[kSnowChunk];
#pragma unroll
    for (int j = 0; j < kSnowChunk; j++) {
      sa[j] = snow_any ? sp[int64_t(j) * n] : WgD2{0., 0.};
      ea[j] = ep[int64_t(j) * n];
    }
#pragma unroll 1
    for (int c = 0; c < WG_DEV_NPAIR / kSnowChunk; c++) {
      const bool more = c + 1 < WG_DEV_NPAIR / kSnowChunk;
      if (more) {
#pragma unroll
        for (int j = 0; j < kSnowChunk; j++) {
          sb[j] = snow_any ? sp[int64_t((c + 1) * kSnowChunk + j) * n] : WgD2{0., 0.};
          eb[j] = ep[int64_t((c + 1) * kSnowChunk + j) * n];
        }
      }
      // Sub-cells without snow above freezing keep a zero storage and add
      // eff_bare = prec_to_soil + 0 to the effective precipitation, the full step's
      // result for them.  The wavefront takes that path for the whole chunk when all
      // its lanes may; otherwise every lane runs the full step, one straight-line
      // block for the chunk, whose sub-cells the compiler interleaves.
      double tx[kSnowChunk], ty[kSnowChunk];
      int ex[kSnowChunk], ey[kSnowChunk];
      bool bare = true;
#pragma unroll
      for (int j = 0; j < kSnowChunk; j++) {
        const int32_t e2 = ea[j];
        ex[j] = int(int16_t(e2 & 0xffff));
        ey[j] = int(int16_t(e2 >> 16));
        tx[j] = sd.tempC - ((ex[j] - sd.elev0) * sd.gradnt);
        ty[j] = sd.tempC - ((ey[j] - sd.elev0) * sd.gradnt);
        bare = bare && sa[j].x == 0. && sa[j].y == 0. && tx[j] > sd.snowfz && ty[j] > sd.snowfz;
      }
      bare = bare && sd.melt_ok;
      if (__builtin_amdgcn_ballot_w64(!bare) == 0) {
        if (c == 0) acc.temp_max = tx[0];
#pragma unroll
        for (int j = 0; j < kSnowChunk; j++) {
          acc.eff += sd.eff_bare;
          acc.eff += sd.eff_bare;
          sa[j] = WgD2{0., 0.};
        }
      } else {
#pragma unroll
        for (int j = 0; j < kSnowChunk; j++) {
          sa[j].x = snow_subcell(sa[j].x, tx[j], ex[j], c == 0 && j == 0, sd, acc);
          sa[j].y = snow_subcell(sa[j].y, ty[j], ey[j], false, sd, acc);
        }
      }
#pragma unroll
      for (int j = 0; j < kSnowChunk; j++) {
        const WgD2 v = sa[j];
        const bool nz = v.x != 0. || v.y != 0.;
        any_after |= nz;
        if (snow_any || nz) sp[int64_t(c * kSnowChunk + j) * n] = v;
      }
      if (more) {
#pragma unroll
        for (int j = 0; j < kSnowChunk; j++) {
          sa[j] = sb[j];
          ea[j] = eb[j];
        }
      }
    }
    X.snow_any[mp] = any_after ? 1 : 0;
    dailySnowEvapo = acc.evapo;
    acc.snow_sum /= 100.;
    acc.eff /= 100.;
    dailySnowEvapo /= 100.;
    acc.change /= 100.;
    landStorageChangeSum += acc.change;
  }
  load_cell_in_b(S, X, mp, p, ci);
  const double snow_sum = acc.snow_sum;
  const double TempElevMax = acc.temp_max;
  dailyEffPrec = acc.eff;
  X.snow[mp] = snow_sum;

  // immediate runoff, daily.cpp:1066-1071
  const float built_up = ci.s_built_up;
  if (built_up > 0.) {
    immediate_runoff = kRunoffFracBuiltUp * dailyEffPrec * built_up;
    dailyEffPrec -= immediate_runoff;
  }

  // soil, runoff, AET, recharge, daily.cpp:1080-1239
  const double cfa = ci.x_cfa;
  const float smax = ci.x_smax;
  double soil = ci.x_soil;
  if (landAreaFrac <= 0.) {
    transfer += soil;
    transfer *= cfa;
    soil = 0.;
    daily_gw_recharge = 0.;
    total_daily_runoff = 0.;
  } else {
    soil *= lafPrev / landAreaFrac;
    initialStorage = soil;
    soil_water_overflow = 0;
    if (soil > smax) {
      soil_water_overflow = soil - smax;
      soil = smax;
    }
    if (TempElevMax > T_SNOWFZ) {
      if (smax > 0.) {
        const double soil_saturation = soil / smax;
        daily_runoff = dailyEffPrec * wg_pow(soil_saturation, ci.x_gamma);
        if (dailySoilPET > (maxDailyPET - dailyCanopyEvapo) * soil_saturation)
          dailyAET = (maxDailyPET - dailyCanopyEvapo) * soil_saturation;
        else
          dailyAET = dailySoilPET;
        soil += dailyEffPrec - dailyAET - daily_runoff;
        if (fabs(soil) <= kMinStorVol) soil = 0.;
        dailyEffPrec = 0.;
        if (soil < 0.) {
          dailyAET += soil;
          soil = 0.;
        }
        daily_runoff *= cfa;
        immediate_runoff *= cfa;
        const double rgmax = ci.s_rgmax / 100.;
        const float fg = ci.s_gw_factor;
        if (rgmax < (fg * daily_runoff))
          daily_gw_recharge = rgmax;
        else
          daily_gw_recharge = fg * daily_runoff;
        pot_gw_recharge = 0.;
        if (((arid_gw) && (ci.s_texture < 21)) && (ci.s_ldd >= 0)) {
          if (dailyPrec <= ci.s_pcritgwa) {
            pot_gw_recharge = daily_gw_recharge;
            daily_gw_recharge = 0.;
          }
        }
        daily_runoff -= pot_gw_recharge;
        pot_gw_recharge /= cfa;
        soil += pot_gw_recharge;
        if (soil > smax) {
          soil_water_overflow += soil - smax;
          soil = smax;
        }
        soil_water_overflow *= cfa;
        total_daily_runoff = daily_runoff + immediate_runoff + soil_water_overflow;
      } else {
        total_daily_runoff = 0.;
        daily_gw_recharge = 0.;
      }
    } else {
      soil_water_overflow *= cfa;
      dailyEffPrec *= cfa;
      total_daily_runoff += soil_water_overflow + dailyEffPrec;
      daily_gw_recharge = 0.;
      dailyAET = 0.;
    }
    landStorageChangeSum += soil - initialStorage;
    double cell_aet = landStorageChangeSum * (cfa - 1.0) - dailyPrec * (cfa - 1.0) +
                      (dailyAET + dailyCanopyEvapo + dailySnowEvapo) * cfa;
    if (cell_aet < 0.) neg_land_aet = cell_aet;
    dailyCellLandAET_uncorr = (dailyAET + dailyCanopyEvapo + dailySnowEvapo);
  }
  // G_dailyGwRecharge is handed to routing before the corrections below (daily.cpp:1214)
  const double recharge_handoff = daily_gw_recharge;
  // surface runoff, daily.cpp:1241-1257
  if (neg_land_aet < 0.) total_daily_runoff = total_daily_runoff + neg_land_aet;
  if (total_daily_runoff < 0.) total_daily_runoff = 0.;
  if ((total_daily_runoff - daily_gw_recharge) < 0.) {
    const double neg_runoff = total_daily_runoff - daily_gw_recharge;
    daily_gw_recharge = total_daily_runoff;
    soil += neg_runoff;
  }
  const double surf_runoff = total_daily_runoff - daily_gw_recharge;
  X.soil[mp] = soil;
  X.transfer[mp] = transfer;
  if (write_diag) {
    X.d_surf[mp] = surf_runoff;
    X.d_recharge[mp] = recharge_handoff;
    X.d_ow_pet[mp] = dailyOpenWaterPET;
    X.d_aet[mp] = dailyCellLandAET_uncorr;
    X.d_prec[mp] = dailyPrec;
  }

  // ------------------------------------------------------------------
  // routing without upstream water, routing.cpp:1877-2617
  // ------------------------------------------------------------------
  const double cellArea = ci.s_cell_area;
  const double cf = ci.s_contfreq;
  const int ldd = ci.s_ldd;
  const double laf = landAreaFrac;      // G_landAreaFrac of the day
  const double laf_prev = pd.laf_prev;
  const double ow_prec = dailyPrec, ow_pet = dailyOpenWaterPET;
  double dailyLocalSurfaceRunoff, fswb_catchment = 0.;
  double localRunoff = 0., lrir = 0., lgwrir = 0., recharge_km3 = 0.;
  if (laf <= 0.)
    dailyLocalSurfaceRunoff = transfer * cellArea / 1000000. * laf_prev / 100.;
  else
    dailyLocalSurfaceRunoff = surf_runoff * cellArea / 1000000. * laf / 100.;
  if (ldd >= 0) {
    fswb_catchment = ci.s_fswb_init * 20.;
    if (fswb_catchment > 1.) fswb_catchment = 1.;
    lrir = (1. - fswb_catchment) * dailyLocalSurfaceRunoff;
  }
  const bool arid = (ci.s_aindex == 1);
  if (arid && (ldd >= 0)) {
    localRunoff = fswb_catchment * dailyLocalSurfaceRunoff;
    recharge_km3 = recharge_handoff * cellArea * (laf / 100.) / 1000000.;
  } else {
    // humid cells (routing.cpp:1968-2039) and inland sinks (2112-2159)
    const double gw_prev = ci.x_gw;
    double netGWin = recharge_handoff * cellArea * (laf / 100.) / 1000000.;
    if (kUse) netGWin -= use_nag;  // routing.cpp:1990-1996, 2131-2137
    double gw = gw_prev * ci.s_e_gw + ci.s_inv_kgw * netGWin * ci.s_ome_gw;
    if (fabs(gw) <= kMinStorVol) gw = 0.;
    double gw_runoff = gw_prev - gw + netGWin;
    if (gw_runoff <= 0.) {
      gw_runoff = 0.;
      gw = gw_prev + netGWin;
      if (fabs(gw) <= kMinStorVol) gw = 0.;
    }
    X.gw[mp] = gw;
    if (ldd >= 0) {
      lgwrir = (1. - fswb_catchment) * gw_runoff;
      const double localGWRunoff = fswb_catchment * gw_runoff;
      localRunoff = (fswb_catchment * dailyLocalSurfaceRunoff) + localGWRunoff;
    } else {
      if (laf == 0.)
        dailyLocalSurfaceRunoff = transfer * cellArea / 1000000. * laf_prev / 100.;
      else
        dailyLocalSurfaceRunoff = surf_runoff * cellArea / 1000000. * laf / 100.;
      localRunoff = dailyLocalSurfaceRunoff + gw_runoff;
    }
  }

  if (X.pot_base) {
    // terms of the potential cell runoff: the local runoff terms (routing.cpp:2045-2047
    // humid, 2180 inland sinks; (semi-)arid cells add their groundwater runoff in the
    // chain, 3372-3373) and the open-water precipitation minus evaporation of
    // G_lakeBalance (daily.cpp:773), which wg_pot_runoff_kernel completes
    X.pot_base[mp] = (ldd < 0) ? localRunoff : (arid ? localRunoff + lrir : (localRunoff + lrir) + lgwrir);
    X.pot_ow[mp] = ow_prec - ow_pet;
  }

  double inflow = localRunoff, outflow = 0.;
  const double k_sw = ci.s_k_sw;
  const double evarexp = ci.s_evarexp;
  const bool arid_swb = arid && (ldd >= 0);
  double gwr_loclak = 0., gwr_locwet = 0.;
  // local lake, routing.cpp:2316-2490
  const double loc_lake = ci.s_loc_lake;
  if (loc_lake > 0.) {
    const double prev = ci.x_loc_lake, rf = pd.red_ll;
    const double maxStorage = ((loc_lake) / 100.) * cellArea * ci.s_lake_depth;
    double evap = ((1.0 - cfa) * ow_prec * rf) + (cfa * (ow_pet * rf));
    if (evap < 0.) evap = 0.;
    const double totalInflow = inflow + (ow_prec * rf) * (cellArea / 1000000.) * (loc_lake / 100.);
    if (arid_swb) gwr_loclak = kKswbgw * rf * loc_lake / 100. / (cf / 100.);
    const double PETgwr = evap * (cellArea / 1000000.) * (loc_lake / 100.) + gwr_loclak * cellArea * (cf / 100.) / 1000000.;
    double PETgwrMax = prev + maxStorage + totalInflow;
    if (PETgwrMax < 0.) PETgwrMax = 0.;
    double st;
    if (PETgwr > PETgwrMax) {
      st = (-1.) * maxStorage;
      gwr_loclak *= PETgwrMax / PETgwr;
    } else {
      st = prev + totalInflow - PETgwr;
    }
    if (prev > 0.) {
      outflow = k_sw * prev * wg_pow((prev / maxStorage), kLakeOutflowExp);
      if (st <= 0.) {
        outflow = 0;
      } else {
        if (outflow > st) outflow = st;
      }
    } else {
      outflow = 0.;
    }
    st -= outflow;
    if (fabs(st) <= kMinStorVol) st = 0.;
    if (st > maxStorage) {
      outflow += (st - maxStorage);
      st = maxStorage;
    }
    inflow = outflow;
    double r = 1. - wg_pow(fabs(st - maxStorage) / (2. * maxStorage), (evarexp * kEvapoRedExp));
    if (r < 0.) r = 0.;
    if (r > 1.) r = 1.;
    X.loc_lake[mp] = st;
    X.red_ll[mp] = r;
  }
  // local wetland, routing.cpp:2495-2617
  const double loc_wetland = ci.s_loc_wetland;
  if (loc_wetland > 0.) {
    const double prev = ci.x_loc_wetl, rf = pd.red_lw;
    const double maxStorage = ((loc_wetland) / 100.) * cellArea * ci.s_wetl_depth;
    double evap = ((1.0 - cfa) * ow_prec * rf) + (cfa * (ow_pet * rf));
    if (evap < 0.) evap = 0.;
    const double totalInflow = inflow + (ow_prec * rf * (cellArea / 1000000.) * (loc_wetland / 100.));
    if (arid_swb) gwr_locwet = kKswbgw * rf * loc_wetland / 100. / (cf / 100.);
    const double PETgwr = evap * (cellArea / 1000000.) * (loc_wetland / 100.) + gwr_locwet * cellArea * (cf / 100.) / 1000000.;
    const double PETgwrMax = prev + totalInflow;
    double st;
    if (PETgwr > PETgwrMax) {
      st = 0.;
      gwr_locwet *= PETgwrMax / PETgwr;
    } else {
      st = prev + totalInflow - PETgwr;
    }
    if (fabs(st) <= kMinStorVol) st = 0.;
    if (st > 0.) {
      outflow = k_sw * st * wg_pow((st / maxStorage), kWetlOutflowExp);
      if (outflow > st) outflow = st;
    } else {
      outflow = 0.;
    }
    st -= outflow;
    if (st > maxStorage) {
      outflow += (st - maxStorage);
      st = maxStorage;
    }
    inflow = outflow;
    double r = 1. - wg_pow(fabs(st - maxStorage) / (maxStorage), (evarexp * kEvapoRedExp));
    if (r < 0.) r = 0.;
    if (r > 1.) r = 1.;
    X.loc_wetl[mp] = st;
    X.red_lw[mp] = r;
  }

  // record of the routing terms that do not depend on the inflow from upstream
  // (routing.cpp:2630-3514)
  const WgChainLink lk = ci.link;
  ChainRec c;
  double *v = c.v;
  const double e_sw = ci.s_e_sw;
  {
    const WgChainStatic &cs = ci.cstat;
    c.inv_ksw = cs.inv_ksw;
    c.ome_sw = cs.ome_sw;
    c.inv_kgw = cs.inv_kgw;
    c.ome_gw = cs.ome_gw;
    c.area = cs.area;
    c.cf100 = cs.cf100;
    c.flags = lk.flags;
  }
  {
    // river velocity and storage terms, routing.cpp:3424-3478
    const double L = ci.s_river_length;
    const double prev = ci.x_river;
    const double velocity = new_river_velocity(ci.s_vel_c1, ci.s_vel_c2, ci.s_bottom_width, prev, L);
    const double K = velocity / L;
    const double eK = wg_exp(-1. * K);
    const double riverAreaFrac = pd.river_frac_next;
    v[CF_INFLOW] = inflow;
    v[CF_LRIR] = lrir;
    v[CF_LGWRIR] = lgwrir;
    v[CF_RV_PREV] = prev;
    v[CF_RV_S0E] = prev * eK;
    v[CF_RV_INVK] = 1. / K;
    v[CF_RV_OME] = 1. - eK;
    v[CF_RV_TERM] = (K * prev * eK) / (1. - eK);
    v[CF_RV_EVAP] = ((1.0 - cfa) * (ow_prec) + (cfa * ow_pet)) * riverAreaFrac / 100. * cellArea / 1000000.;
    v[CF_RV_PREC] = ow_prec * riverAreaFrac / 100. * cellArea / 1000000.;
  }
  v[CF_OW_PREC] = ow_prec;
  v[CF_OW_PET] = ow_pet;
  {
    // global lake, routing.cpp:2633-2676
    const double lake_area = ci.s_lake_area;
    v[CF_GL_LP] = 0.;
    v[CF_GL_USE] = 0.;
    v[CF_GL_MAX] = 0.;
    v[CF_GL_PREV] = 0.;
    v[CF_GL_S0E] = 0.;
    v[CF_GL_GWR] = 0.;
    if (lake_area > 0.) {
      const double prev = ci.x_glo_lake, rf = pd.red_gl;
      double evap = ((1.0 - cfa) * ow_prec) + (cfa * ow_pet * rf);
      if (evap < 0.) evap = 0.;
      const double gwr = arid_swb ? kKswbgw * rf * (lake_area / (cellArea * (cf / 100.))) : 0.;
      v[CF_GL_MAX] = lake_area * ci.s_lake_depth;
      v[CF_GL_LP] = (ow_prec * (lake_area / 1000000.));
      v[CF_GL_GWR] = gwr;
      // the lake's share of the use, half of it when the cell has a reservoir too
      // (routing.cpp:2680-2695)
      const double use_lake = (lk.flags & CFL_RES) ? 0.5 * use_desired : use_desired;
      v[CF_GL_USE] = evap * (lake_area / 1000000.) + gwr * cellArea * (cf / 100.) / 1000000. +
                     (kUse ? use_lake : 0.);
      v[CF_GL_PREV] = prev;
      v[CF_GL_S0E] = prev * e_sw;
    }
  }
  {
    // global wetland, routing.cpp:3181-3222
    const double glo_wetland = ci.s_glo_wetland;
    v[CF_GW_WP] = 0.;
    v[CF_GW_PET] = 0.;
    v[CF_GW_MAX] = 0.;
    v[CF_GW_PREV] = 0.;
    v[CF_GW_S0E] = 0.;
    v[CF_GW_GWR] = 0.;
    if (glo_wetland > 0) {
      const double prev = ci.x_glo_wetl, rf = pd.red_gw;
      double evap = ((1.0 - cfa) * (ow_prec * rf)) + (cfa * (ow_pet * rf));
      if (evap < 0.) evap = 0.;
      const double gwr = arid_swb ? kKswbgw * rf * glo_wetland / 100. / (cf / 100.) : 0.;
      v[CF_GW_MAX] = ((glo_wetland) / 100.) * cellArea * ci.s_wetl_depth;
      v[CF_GW_WP] = (ow_prec * rf * (cellArea / 1000000.) * (glo_wetland / 100.));
      v[CF_GW_GWR] = gwr;
      v[CF_GW_PET] = evap * (cellArea / 1000000.) * ((glo_wetland) / 100.) + gwr * cellArea * (cf / 100.) / 1000000.;
      v[CF_GW_PREV] = prev;
      v[CF_GW_S0E] = prev * e_sw;
    }
  }
  {
    // groundwater of (semi-)arid cells, routing.cpp:3310-3330
    v[CF_A_GWR_LL] = gwr_loclak;
    v[CF_A_GWR_LW] = gwr_locwet;
    v[CF_A_RECH] = recharge_km3;
    const double gprev = arid_swb ? ci.x_gw : 0.;
    v[CF_A_PREV] = gprev;
    v[CF_A_S0E] = gprev * ci.s_e_gw;
  }
  {
    // reservoir, routing.cpp:2807-2860 inputs
    const bool res = (lk.flags & CFL_RES) != 0;
    v[CF_RES_PREV] = res ? ci.x_res : 0.;
    v[CF_RES_RED] = res ? pd.red_res : 0.;
    v[CF_RES_CFA] = cfa;
    v[CF_RES_KREL] = res ? ci.x_k_release : 0.;
    v[CF_RES_CAP] = res ? ci.s_stor_cap : 0.;
    v[CF_RES_QMEAN] = res ? ci.s_mean_outflow : 0.;
    v[CF_RES_AREA] = res ? ci.s_reservoir_area : 0.;
    v[CF_RES_PROVREL] = res ? (kUse ? S.use_provrel[use_t] : ci.s_prov_rel) : 0.;
    v[CF_RES_PAD] = 0.;
  }
  v[CF_USE] = use_desired;
  v[CF_A_NUG] = use_nag;
  if (lk.nu == 0) {
    // no upstream cell: G_riverInflow[n] is still 0 when the cell is routed
    const ChainOut o = chain_step<kUse>(c, 0.0, day_info(doy));
    store_chain(X, mp, c.flags, o);
    if (lk.station >= 0)
      X.st_daily[(int64_t(m) * WG_DAILY_SLOTS + daily_slot) * S.n_stations + lk.station] = (ldd >= 0) ? o.tv : 0.;
  } else {
    // the prefix for every routed cell, the suffix for cells with a global lake or reservoir
    // the pairs holding the prefix for every routed cell (the last of them also
    // holds the first suffix field), the rest for cells with a lake or reservoir
    WgD2 *dst = rec_at(X, n, m, p);
    constexpr int kPrefixPairs = (CF_PREFIX + 1) / 2;
#pragma unroll
    for (int k = 0; k < kPrefixPairs; k++) dst[int64_t(k) * n] = WgD2{v[2 * k], v[2 * k + 1]};
    if (lk.flags & (CFL_LAKE | CFL_RES)) {
#pragma unroll
      for (int k = kPrefixPairs; k < CF_USE / 2; k++) dst[int64_t(k) * n] = WgD2{v[2 * k], v[2 * k + 1]};
    }
    if (kUse) dst[int64_t(CF_USE / 2) * n] = WgD2{v[CF_USE], v[CF_A_NUG]};
  }
}

// Workgroup barrier that orders LDS only.  The level walk exchanges data between
// threads through LDS alone (its global stores are read by later kernels), and a
// full __syncthreads() would also wait for the next level's prefetch loads
// (vmcnt counts loads and stores alike on this architecture).
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Inputs of one chain cell, loaded a level (or two) before it is routed: the
// whole routing record, the outflow constants and the links.  The record suffix
// holds data only for cells with a global lake or reservoir; it is loaded for
// every cell (so that no load depends on the flags) and zeroed for the others.
struct ChainSlot {
  double a[CF_COUNT];
  double inv_ksw, ome_sw, inv_kgw, ome_gw, area, cf100;
  int flags, nu, station, level, nu_level;
  int up[4];  // group-local upstream indices, two int16 per word
};

__device__ inline void prefetch_slot(const WgStatic &S, const WgState &X, int64_t base, int p, ChainSlot &q) {
  const int64_t n = S.n;
  const WgChainStatic cs = load_cstat(S, p);
  const WgD2 *src = X.crec + base * (WG_DEV_NREC / 2) + p;  // rec_at(X, n, m, p) for base = m * n
#pragma unroll
  for (int k = 0; k < CF_COUNT / 2; k++) {
    const WgD2 v = src[k * n];
    q.a[2 * k] = v.x;
    q.a[2 * k + 1] = v.y;
  }
  q.inv_ksw = cs.inv_ksw;
  q.ome_sw = cs.ome_sw;
  q.inv_kgw = cs.inv_kgw;
  q.ome_gw = cs.ome_gw;
  q.area = cs.area;
  q.cf100 = cs.cf100;
  const WgI4 a = S.clink[p], b = S.clink[n + p];
  q.flags = a.x;
  q.station = a.y;
  q.nu = a.z & 0xffff;
  q.nu_level = a.z >> 16;
  q.level = a.w;
  q.up[0] = b.x;
  q.up[1] = b.y;
  q.up[2] = b.z;
  q.up[3] = b.w;
}

typedef unsigned int WgU32x4 __attribute__((ext_vector_type(4)));

// Wave-walk variant of prefetch_slot: the record suffix is fetched only for cells
// with a global lake or reservoir (about 1.5 % of routed cells).  The cell's chain
// flags `fl` come from LDS, so no global load waits for another.  The other lanes
// issue the same suffix loads through raw buffers at an offset past the row's end,
// which the hardware answers with zeros without a memory access: every lane issues
// the same loads whatever its flags, so the count of outstanding loads, which the
// compiler's waits rely on, does not depend on them.
template <bool kUse>
__device__ inline void prefetch_slot_wave(const WgStatic &S, const WgState &X, int64_t base, int p, int fl,
                                          ChainSlot &q) {
  const int64_t n = S.n;
  const WgChainStatic cs = load_cstat(S, p);
  const WgD2 *src = X.crec + base * (WG_DEV_NREC / 2) + p;
  constexpr int kPrefixPairs = (CF_PREFIX + 1) / 2;
#pragma unroll
  for (int k = 0; k < kPrefixPairs; k++) {
    const WgD2 v = src[k * n];
    q.a[2 * k] = v.x;
    q.a[2 * k + 1] = v.y;
  }
  const bool suffix = (fl & (CFL_LAKE | CFL_RES)) != 0;
  const int off = suffix ? p * 16 : int(0x80000000u);
  WgD2 *row0 = X.crec + base * (WG_DEV_NREC / 2);
#pragma unroll
  for (int k = kPrefixPairs; k < CF_USE / 2; k++) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(row0 + k * n, (short)0, int(n * 16), 0x00020000);
    const WgU32x4 w = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    q.a[2 * k] = __builtin_bit_cast(double, (unsigned long long)w.x | ((unsigned long long)w.y << 32));
    q.a[2 * k + 1] = __builtin_bit_cast(double, (unsigned long long)w.z | ((unsigned long long)w.w << 32));
  }
  // the water-use pair, for every cell when water use is on
  if (kUse) {
    const WgD2 u = src[int64_t(CF_USE / 2) * n];
    q.a[CF_USE] = u.x;
    q.a[CF_A_NUG] = u.y;
  } else {
    q.a[CF_USE] = 0.;
    q.a[CF_A_NUG] = 0.;
  }
  q.inv_ksw = cs.inv_ksw;
  q.ome_sw = cs.ome_sw;
  q.inv_kgw = cs.inv_kgw;
  q.ome_gw = cs.ome_gw;
  q.area = cs.area;
  q.cf100 = cs.cf100;
  const WgI4 a = S.clink[p], b = S.clink[n + p];
  q.flags = a.x;
  q.station = a.y;
  q.nu = a.z & 0xffff;
  q.nu_level = a.z >> 16;
  q.level = a.w;
  q.up[0] = b.x;
  q.up[1] = b.y;
  q.up[2] = b.z;
  q.up[3] = b.w;
}

// Wide-level variant: a batch of one level takes the record suffix when one of
// its cells has a global lake or reservoir (`suffix`, uniform over the wavefront,
// from the batch table), and issues no suffix loads otherwise.  Lake and reservoir
// cells come last in their level, so most batches load the prefix alone.  The
// workgroup's wavefronts share one CU's vector memory pipeline, which bounds the
// wide levels: every load instruction left out shortens them.
template <bool kUse>
__device__ inline void prefetch_slot_batch(const WgStatic &S, const WgState &X, int64_t base, int p, bool suffix,
                                           ChainSlot &q) {
  const int64_t n = S.n;
  const WgChainStatic cs = load_cstat(S, p);
  const WgD2 *src = X.crec + base * (WG_DEV_NREC / 2) + p;
  constexpr int kPrefixPairs = (CF_PREFIX + 1) / 2;
#pragma unroll
  for (int k = 0; k < kPrefixPairs; k++) {
    const WgD2 v = src[k * n];
    q.a[2 * k] = v.x;
    q.a[2 * k + 1] = v.y;
  }
  if (suffix) {
#pragma unroll
    for (int k = kPrefixPairs; k < CF_USE / 2; k++) {
      const WgD2 v = src[k * n];
      q.a[2 * k] = v.x;
      q.a[2 * k + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int k = 2 * kPrefixPairs; k < CF_USE; k++) q.a[k] = 0.;
  }
  if (kUse) {
    const WgD2 u = src[int64_t(CF_USE / 2) * n];
    q.a[CF_USE] = u.x;
    q.a[CF_A_NUG] = u.y;
  } else {
    q.a[CF_USE] = 0.;
    q.a[CF_A_NUG] = 0.;
  }
  q.inv_ksw = cs.inv_ksw;
  q.ome_sw = cs.ome_sw;
  q.inv_kgw = cs.inv_kgw;
  q.ome_gw = cs.ome_gw;
  q.area = cs.area;
  q.cf100 = cs.cf100;
  const WgI4 a = S.clink[p], b = S.clink[n + p];
  q.flags = a.x;
  q.station = a.y;
  q.nu = a.z & 0xffff;
  q.nu_level = a.z >> 16;
  q.level = a.w;
  q.up[0] = b.x;
  q.up[1] = b.y;
  q.up[2] = b.z;
  q.up[3] = b.w;
}

// The routing record of a slot (the suffix of a cell without a global lake or
// reservoir holds no data; chain_step does not read it for such cells).
__device__ inline void slot_record(const ChainSlot &q, ChainRec &c) {
#pragma unroll
  for (int f = 0; f < CF_COUNT; f++) c.v[f] = q.a[f];
  c.inv_ksw = q.inv_ksw;
  c.ome_sw = q.ome_sw;
  c.inv_kgw = q.inv_kgw;
  c.ome_gw = q.ome_gw;
  c.area = q.area;
  c.cf100 = q.cf100;
  c.flags = q.flags;
}

// Stores of the level walk.  Every lane issues the same stores on every level,
// whether or not it has a cell there and whatever storages the cell has: a store
// that must not land goes through a raw buffer over the destination row at an
// offset past the row's end, which the hardware drops.  A store behind a branch
// would leave different counts of outstanding memory operations on the two sides
// of the branch; the compiler's wait for the next level's inputs must then assume
// the smaller count, and waits for the loads just issued for the level after
// (loads and stores share one in-order counter on this architecture).
typedef unsigned int WgU32x2 __attribute__((ext_vector_type(2)));
// s_waitcnt immediate of the gfx9 family for vmcnt(0) (expcnt and lgkmcnt at their
// maxima, i.e. not waited on)
constexpr int kWaitVmcnt0 = 0x0F70;
constexpr uint32_t kDropOffset = 0x80000000u;  // past every row (rows hold < 2^28 doubles)

__device__ inline __amdgpu_buffer_rsrc_t row_buffer(double *row, int n) {
  // raw buffer (stride 0) of n doubles; the descriptor's last word is the one the
  // gfx9 family takes for untyped 32-bit dword access
  return __builtin_amdgcn_make_buffer_rsrc(row, (short)0, n * 8, 0x00020000);
}

__device__ inline void store_or_drop(double *row, int n, bool keep, int i, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(WgU32x2, v), row_buffer(row, n),
                                        int(keep ? uint32_t(i) * 8u : kDropOffset), 0, 0);
}

// One chain cell: ordered gather of the upstream transported volumes (the serial
// += of routing.cpp:3955-3958, in routing order), the upstream-dependent routing
// step, LDS and global stores, the station's daily discharge.  `act` is false for
// a lane without a cell on this level (its slot holds some other cell of the
// group, its results are not stored).
template <bool kUse>
__device__ __forceinline__ void chain_cell(const WgStatic &S, const WgState &X, double *s_tv, const ChainSlot &q,
                                           bool act, int64_t base, int p, int gs, int m, const DayInfo &d,
                                           int daily_slot) {
  const bool run = act && (q.flags & CFL_CONT);
  double up = 0.;
#pragma unroll
  for (int u = 0; u < 8; u++)
    if (u < q.nu) up += s_tv[(q.up[u >> 1] >> ((u & 1) * 16)) & 0xffff];
  ChainRec c;
  slot_record(q, c);
  const ChainOut o = chain_step<kUse>(c, up, d);
  const int n = S.n;
  store_or_drop(X.glo_lake + base, n, run && (q.flags & CFL_LAKE), p, o.glo_lake);
  store_or_drop(X.res + base, n, run && (q.flags & CFL_RES), p, o.res);
  store_or_drop(X.k_release + base, n, run && (q.flags & CFL_RES), p, o.k_release);
  store_or_drop(X.glo_wetl + base, n, run && (q.flags & CFL_GWET), p, o.glo_wetl);
  store_or_drop(X.gw + base, n, run && (q.flags & CFL_ARID), p, o.gw);
  if (X.pot_lgw) store_or_drop(X.pot_lgw + base, n, run && (q.flags & CFL_ARID), p, o.lgw);
  if (kUse) store_or_drop(X.use_rem + base, n, run, p, o.rem);
  store_or_drop(X.river + base, n, run, p, o.river);
  store_or_drop(X.tv + base, n, run, p, o.tv);
  if (run) s_tv[p - gs] = o.tv;
  const double qq = (run && (q.flags & CFL_LDD)) ? o.tv : 0.;
  store_or_drop(X.st_daily + (int64_t(m) * WG_DAILY_SLOTS + daily_slot) * S.n_stations, S.n_stations,
                act && q.station >= 0, q.station, qq);
}

// The first cell of level lf at this thread's position, or the group's first cell
// when there is none (or lf is past the group): the slot is loaded either way.
__device__ inline int slot_cell(const int32_t *__restrict__ lo, int lf, int l1, int gs) {
  const int lc = min(lf, l1 - 1);
  const int p = lo[lc] + int(threadIdx.x);
  return (lf < l1 && p < lo[lc + 1]) ? p : gs;
}

// One level of a group: the thread's first cell comes from `use` (loaded kAhead
// levels earlier); the first cell of level l + kAhead is loaded into `fill`
// before any computation, so its latency overlaps the levels in between.  The
// slots rotate by name, never copied, so no wait on the new loads is forced.
// Cells beyond the block size (levels wider than the block) load directly.
template <int kAhead, bool kUse>
__device__ __forceinline__ void chain_level(const WgStatic &S, const WgState &X, double *s_tv,
                                            const int32_t *__restrict__ lo, int l, int l1, int64_t base, int gs,
                                            int m, const DayInfo &d, int daily_slot, const ChainSlot &use,
                                            ChainSlot &fill) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int beg = lo[l], end = lo[l + 1];
  prefetch_slot(S, X, base, slot_cell(lo, l + kAhead, l1, gs), fill);
  const bool act = beg + tid < end;
  chain_cell<kUse>(S, X, s_tv, use, act, base, act ? beg + tid : gs, gs, m, d, daily_slot);
  for (int p = beg + tid + nt; p < end; p += nt) {
    ChainSlot q;
    prefetch_slot(S, X, base, p, q);
    chain_cell<kUse>(S, X, s_tv, q, true, base, p, gs, m, d, daily_slot);
  }
  lds_barrier();
}

// One workgroup per (routing group, member): levels 1..depth of the group.  The
// transported volumes of the group live in LDS (level 0 is copied in from the
// cell kernel's output); the static record holds group-local upstream indices.
// Each thread's first cell of a level is loaded kAhead levels early (1 or 2), so
// that the load latency overlaps the levels in between; the slots rotate by name.
template <int kAhead, bool kUse>
__global__ __launch_bounds__(256) void wg_chain_kernel(WgStatic S, WgState X, WgTables T,
                                                       const int32_t *__restrict__ group_level_off,
                                                       const int32_t *__restrict__ group_first_level,
                                                       const WgClock *__restrict__ clock, int k,
                                                       int daily_slot) {
  extern __shared__ double s_tv[];
  const int g = S.block_group[blockIdx.x];
  if (g < 0) return;
  const int m = blockIdx.y;
  const int n = S.n;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int doy = int((clock->doy0 - 1 + k) % 365) + 1;
  const DayInfo d = day_info(doy);
  const int32_t *lo = group_level_off;
  const int l0 = group_first_level[g], l1 = group_first_level[g + 1];
  const int gs = lo[l0], ge = lo[l1];
  if (l1 - l0 < 2 || ge <= gs) return;  // nothing routed in this kernel
  const int lvl0_end = lo[l0 + 1];
  const int64_t base = int64_t(m) * n;
  ChainSlot a, b, c;
  prefetch_slot(S, X, base, slot_cell(lo, l0 + 1, l1, gs), a);
  if (kAhead == 2) prefetch_slot(S, X, base, slot_cell(lo, l0 + 2, l1, gs), b);
  // level 0 was routed by the cell kernel; cells not routed today read as 0.
  // Eight loads in flight per thread before their LDS stores.
  for (int i0 = gs + tid; i0 < ge; i0 += 8 * nt) {
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int i = i0 + j * nt;
      v[j] = (i < lvl0_end) ? X.tv[base + i] : 0.;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int i = i0 + j * nt;
      if (i < ge) s_tv[i - gs] = v[j];
    }
  }
  // the first slots' loads complete here on every path into the level loop, so
  // that the loop's waits count only what the loop itself issues
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  lds_barrier();
  if (kAhead == 1) {
    for (int l = l0 + 1; l < l1; l += 2) {
      chain_level<1, kUse>(S, X, s_tv, lo, l, l1, base, gs, m, d, daily_slot, a, b);
      if (l + 1 < l1) chain_level<1, kUse>(S, X, s_tv, lo, l + 1, l1, base, gs, m, d, daily_slot, b, a);
    }
  } else {
    for (int l = l0 + 1; l < l1; l += 3) {
      chain_level<2, kUse>(S, X, s_tv, lo, l, l1, base, gs, m, d, daily_slot, a, c);
      if (l + 1 < l1) chain_level<2, kUse>(S, X, s_tv, lo, l + 1, l1, base, gs, m, d, daily_slot, b, a);
      if (l + 2 < l1) chain_level<2, kUse>(S, X, s_tv, lo, l + 2, l1, base, gs, m, d, daily_slot, c, b);
    }
  }
  (void)T;
}

// Wave walk of a routing group.  The levels above level 0 are routed in batches of
// up to 64 cells, one cell per lane, each lane's inputs loaded one batch before
// the batch is routed; transported volumes pass between lanes through LDS.
//  - The leading wide levels (more than 64 cells) are cut into level-aligned
//    batches dealt round-robin to the workgroup's kWaves wavefronts (4 or 8),
//    with an LDS-only barrier after each level.
//  - The narrow rest is routed by wavefront 0 alone in batches of 64 cells taken
//    in device order (by level), so that a batch covers several narrow levels;
//    within a batch the levels run in order, each on the lanes whose cells lie
//    on it.  LDS serves the LDS operations of one wavefront in issue order, so
//    no barrier is needed there.
// A batch's results are stored when the batch is done.  Every batch issues the
// same loads and stores (padding batches route nothing), so that no wait covers
// the loads issued for the next batch.
__device__ inline void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront", "local");
}

// One batch: cells [b0, bend) (b0 >= bend: a padding batch); `q` holds the
// batch's inputs, `fill` receives those of the batch starting at pnext.
// kWide: a batch of the wide levels, where next_lr / cur_lr (uniform) are the
// batches' lakeres flags: the next batch's suffix is loaded, and this batch's
// lake and reservoir stores issued, only when it is 1.  Otherwise a batch of the
// narrow rest: per-cell record suffix through s_fl, every store issued.
template <bool kWide, bool kUse>
__device__ __forceinline__ void wave_batch(const WgStatic &S, const WgState &X, double *s_tv, int b0, int bend,
                                           int pnext, int ge, int gs, int64_t base, int m, const DayInfo &d,
                                           int daily_slot, const uint8_t *s_fl, int next_lr, int cur_lr,
                                           const ChainSlot &q, ChainSlot &fill) {
  const int lane = threadIdx.x & 63;
  {
    const int pf = pnext + lane < ge ? pnext + lane : gs;
    if (kWide)
      prefetch_slot_batch<kUse>(S, X, base, pf, next_lr != 0, fill);
    else
      prefetch_slot_wave<kUse>(S, X, base, pf, s_fl[pf - gs], fill);
  }
  const int p = b0 + lane;
  const bool have = p < bend;
  ChainOut o = {0., 0., 0., 0., 0., 0., 0., 0.};
  bool run = false;
  if (b0 < bend) {
    const int last = min(63, bend - 1 - b0);
    const int lv_a = __builtin_amdgcn_readfirstlane(q.level);
    const int lv_b = __builtin_amdgcn_readlane(q.level, last);
    ChainRec c;
    slot_record(q, c);
    const StepPre pre = step_pre<kUse>(c);
    for (int l = lv_a; l <= lv_b; l++) {
      if (have && q.level == l) {
        // upstream transported volumes in routing order (routing.cpp:3955-3958);
        // entries past the cell's own count read entry 0 and add +0, which leaves
        // the non-negative sum unchanged.  Past the level's largest count (most
        // levels have at most two upstream cells) nothing is added.
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t[u] = s_tv[(q.up[u >> 1] >> ((u & 1) * 16)) & 0xffff];
        double up = 0.;
        up += t[0];
        up += (q.nu > 1) ? t[1] : 0.;
        if (__builtin_amdgcn_readfirstlane(q.nu_level) > 2) {
#pragma unroll
          for (int u = 2; u < 8; u++) up += (u < q.nu) ? t[u] : 0.;
        }
        bool slow;
        o = step_fast<kUse>(c, pre, up, slow);
        if (slow) o = chain_step<kUse>(c, up, d);
        run = (q.flags & CFL_CONT) != 0;
        s_tv[p - gs] = run ? o.tv : 0.;
      }
      wave_lds_order();
    }
  }
  const int n = S.n;
  if (cur_lr != 0) {
    store_or_drop(X.glo_lake + base, n, run && (q.flags & CFL_LAKE), p, o.glo_lake);
    store_or_drop(X.res + base, n, run && (q.flags & CFL_RES), p, o.res);
    store_or_drop(X.k_release + base, n, run && (q.flags & CFL_RES), p, o.k_release);
  }
  store_or_drop(X.glo_wetl + base, n, run && (q.flags & CFL_GWET), p, o.glo_wetl);
  store_or_drop(X.gw + base, n, run && (q.flags & CFL_ARID), p, o.gw);
  if (X.pot_lgw) store_or_drop(X.pot_lgw + base, n, run && (q.flags & CFL_ARID), p, o.lgw);
  if (kUse) store_or_drop(X.use_rem + base, n, run, p, o.rem);
  store_or_drop(X.river + base, n, run, p, o.river);
  store_or_drop(X.tv + base, n, run, p, o.tv);
  const double qq = (run && (q.flags & CFL_LDD)) ? o.tv : 0.;
  store_or_drop(X.st_daily + (int64_t(m) * WG_DAILY_SLOTS + daily_slot) * S.n_stations, S.n_stations,
                have && q.station >= 0, q.station, qq);
}

template <int kWaves, bool kUse>
__global__ __launch_bounds__(64 * kWaves) void wg_chain_wave_kernel(WgStatic S, WgState X, WgTables T,
                                                            const int32_t *__restrict__ group_level_off,
                                                            const int32_t *__restrict__ group_first_level,
                                                            const int32_t *__restrict__ wave_wide,
                                                            const int32_t *__restrict__ wave_item_off,
                                                            const WgWaveItem *__restrict__ wave_items,
                                                            const WgClock *__restrict__ clock, int k,
                                                            int daily_slot) {
  extern __shared__ double s_tv[];
  const int g = S.block_group[blockIdx.x];
  if (g < 0) return;
  const int m = blockIdx.y;
  const int n = S.n;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // the wavefront, uniform
  const int doy = int((clock->doy0 - 1 + k) % 365) + 1;
  const DayInfo d = day_info(doy);
  const int32_t *lo = group_level_off;
  const int l0 = group_first_level[g], l1 = group_first_level[g + 1];
  const int gs = lo[l0], ge = lo[l1];
  if (l1 - l0 < 2 || ge <= gs) return;  // nothing routed in this kernel
  const int lvl0_end = lo[l0 + 1];
  const int64_t base = int64_t(m) * n;
  const int la = wave_wide[g];  // wide levels l0 + 1 .. l0 + la
  const int pb = lo[l0 + 1 + la];  // first cell of the narrow rest
  // level 0 was routed by the cell kernel: its transported volumes, four loads
  // in flight per thread
  constexpr int kThreads = 64 * kWaves;
  // the chain flags of the group's cells, behind its transported volumes (the
  // launch gives every workgroup 9 bytes of LDS per cell of the largest group);
  // both are loaded together, 16 of each in flight per thread
  uint8_t *s_fl = reinterpret_cast<uint8_t *>(s_tv + (ge - gs));
  constexpr int kFill = 16;
  for (int i0 = gs + tid; i0 < ge; i0 += kFill * kThreads) {
    double v[kFill];
    int8_t f[kFill];
#pragma unroll
    for (int j = 0; j < kFill; j++) {
      const int i = i0 + j * kThreads;
      v[j] = (i < lvl0_end) ? X.tv[base + i] : 0.;
      f[j] = (i < ge) ? S.cflags[i] : int8_t(0);
    }
#pragma unroll
    for (int j = 0; j < kFill; j++) {
      const int i = i0 + j * kThreads;
      if (i < lvl0_end) s_tv[i - gs] = v[j];
      if (i < ge) s_fl[i - gs] = uint8_t(f[j]);
    }
  }
  ChainSlot s0, s1;
  // wide levels: this wavefront's batches in level order; a padding batch (past
  // the list) routes nothing and passes no barrier.  Every wavefront passes
  // exactly la barriers, one after each wide level.  The first batch's loads go
  // out with the fill's.
  const int it0 = la > 0 ? wave_item_off[g * kWaves + w] : 0;
  const int it1 = la > 0 ? wave_item_off[g * kWaves + w + 1] : 0;
  auto item = [&](int i) {
    return (i < it1) ? wave_items[i] : WgWaveItem{ge, ge, 0, 0};
  };
  WgWaveItem ia = item(it0);
  if (la > 0) {
    const int pf = ia.start + lane < ge ? ia.start + lane : gs;
    prefetch_slot_batch<kUse>(S, X, base, pf, ia.lakeres != 0, s0);
  }
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  lds_barrier();
  if (la > 0) {
    int lev = 1;
    for (int i = it0; i < it1; i += 2) {
      const WgWaveItem ib = item(i + 1), ic = item(i + 2);
      while (lev < ia.level) {
        lds_barrier();
        lev++;
      }
      wave_batch<true, kUse>(S, X, s_tv, ia.start, ia.end, ib.start, ge, gs, base, m, d, daily_slot, s_fl, ib.lakeres,
                 ia.lakeres, s0, s1);
      while (lev < ib.level) {
        lds_barrier();
        lev++;
      }
      wave_batch<true, kUse>(S, X, s_tv, ib.start, ib.end, ic.start, ge, gs, base, m, d, daily_slot, s_fl, ic.lakeres,
                 ib.lakeres, s1, s0);
      ia = ic;
    }
    while (lev <= la) {
      lds_barrier();
      lev++;
    }
  }
  if (w != 0 || pb >= ge) return;
  // the narrow rest, wavefront 0 alone
  const int nb = (ge - pb + 127) / 128 * 2;
  {
    const int pa = pb + lane < ge ? pb + lane : gs;
    prefetch_slot_wave<kUse>(S, X, base, pa, s_fl[pa - gs], s0);
  }
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  for (int b = 0; b < nb; b += 2) {
    const int b0 = pb + b * 64;
    wave_batch<false, kUse>(S, X, s_tv, b0, ge, b0 + 64, ge, gs, base, m, d, daily_slot, s_fl, 1, 1, s0, s1);
    wave_batch<false, kUse>(S, X, s_tv, b0 + 64, ge, b0 + 128, ge, gs, base, m, d, daily_slot, s_fl, 1, 1, s1, s0);
  }
  (void)T;
}

// Dataflow walk of a routing group, without a barrier per level.  Each thread owns
// the cells of the group's levels at positions tid, tid + nt, ... (so its cells
// come in level order) and routes each one as soon as the transported volumes of
// all its upstream cells are in LDS; LDS holds a sentinel bit pattern until a
// cell's volume is written.  The unfinished cell of lowest level always has its
// inputs ready and its owner at it, so the walk cannot stall; a spin limit that
// every lane reaches ends the kernel (with an error flag) should that ever fail.
constexpr unsigned long long kTvPending = 0xFFF4A11DDEADBEEFull;  // a NaN no arithmetic produces
constexpr uint32_t kFlowSpinLimit = 1u << 24;

__device__ inline bool flow_next(const int32_t *s_lo, int l1_rel, int nt, int &l, int &p) {
  // advance (l, p) to this thread's next cell: same level + nt, else the next level
  // with a cell at this thread's position (l relative to the group's first level)
  p += nt;
  const int tid = threadIdx.x;
  while (l < l1_rel && p >= s_lo[l + 1]) {
    l++;
    if (l < l1_rel) p = s_lo[l] + tid;
  }
  return l < l1_rel;
}

template <bool kUse>
__global__ __launch_bounds__(256) void wg_chain_flow_kernel(WgStatic S, WgState X, WgTables T,
                                                            const int32_t *__restrict__ group_level_off,
                                                            const int32_t *__restrict__ group_first_level,
                                                            const WgClock *__restrict__ clock, int k,
                                                            int daily_slot) {
  extern __shared__ double s_tv[];
  __shared__ int32_t s_lo[WG_MAX_GROUP_DEPTH + 1];
  const int g = S.block_group[blockIdx.x];
  if (g < 0) return;
  const int m = blockIdx.y;
  const int n = S.n;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int doy = int((clock->doy0 - 1 + k) % 365) + 1;
  const DayInfo d = day_info(doy);
  const int l0 = group_first_level[g], l1 = group_first_level[g + 1];
  const int depth = l1 - l0;
  const int gs = group_level_off[l0], ge = group_level_off[l1];
  const int lvl0_end = group_level_off[l0 + 1];
  const int64_t base = int64_t(m) * n;
  for (int i = tid; i <= depth; i += nt) s_lo[i] = group_level_off[l0 + i];
  volatile unsigned long long *vtv = reinterpret_cast<volatile unsigned long long *>(s_tv);
  for (int i = gs + tid; i < ge; i += nt) {
    double v;
    if (i < lvl0_end)
      v = X.tv[base + i];  // level 0 was routed by the cell kernel
    else
      v = (S.clink[i].x & CFL_CONT) ? __longlong_as_double(kTvPending) : 0.;
    s_tv[i - gs] = v;
  }
  __syncthreads();
  int l = 1, p = s_lo[1] + tid;  // level relative to l0
  while (l < depth && p >= s_lo[l + 1]) {
    l++;
    if (l < depth) p = s_lo[l] + tid;
  }
  bool have = l < depth;
  ChainSlot cur, nxt;
  int nl = l, np = p;
  bool have_nxt = false;
  if (have) {
    prefetch_slot(S, X, base, p, cur);
    have_nxt = flow_next(s_lo, depth, nt, nl, np);
    if (have_nxt) prefetch_slot(S, X, base, np, nxt);
  }
  uint32_t spins = 0;
  while (have) {
    if (++spins > kFlowSpinLimit) {
      X.flow_error[0] = 1;
      break;
    }
    bool ready = true;
    if (cur.flags & CFL_CONT) {
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (u < cur.nu && vtv[(cur.up[u >> 1] >> ((u & 1) * 16)) & 0xffff] == kTvPending) ready = false;
    }
    if (!ready) continue;
    double qq = 0.;
    if (cur.flags & CFL_CONT) {
      double up = 0.;
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (u < cur.nu) up += __longlong_as_double(vtv[(cur.up[u >> 1] >> ((u & 1) * 16)) & 0xffff]);
      const int64_t mp = base + p;
      ChainRec c;
      slot_record(cur, c);
      const ChainOut o = chain_step<kUse>(c, up, d);
      vtv[p - gs] = __double_as_longlong(o.tv);
      store_chain(X, mp, cur.flags, o);
      qq = (cur.flags & CFL_LDD) ? o.tv : 0.;
    }
    if (cur.station >= 0)
      X.st_daily[(int64_t(m) * WG_DAILY_SLOTS + daily_slot) * S.n_stations + cur.station] = qq;
    have = have_nxt;
    if (have) {
      cur = nxt;
      l = nl;
      p = np;
      have_nxt = flow_next(s_lo, depth, nt, nl, np);
      if (have_nxt) prefetch_slot(S, X, base, np, nxt);
    }
  }
  (void)T;
}

// Adds the daily station discharge of slots [slot0, slot0 + n_slots) to the running
// sums, day by day in order (the per-day += of routing.cpp:4232-4238).
__global__ void wg_station_reduce(WgState X, int32_t n_stations, int32_t n_members, int slot0, int n_slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_stations * n_members) return;
  const int m = i / n_stations, st = i - m * n_stations;
  double s = X.st_sum[i];
  const double *q = X.st_daily + (int64_t(m) * WG_DAILY_SLOTS + slot0) * n_stations + st;
  int d = 0;
  for (; d + 8 <= n_slots; d += 8) {  // eight loads in flight, added in day order
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = q[int64_t(d + j) * n_stations];
#pragma unroll
    for (int j = 0; j < 8; j++) s += v[j];
  }
  for (; d < n_slots; d++) s += q[int64_t(d) * n_stations];
  X.st_sum[i] = s;
}

// Adds the day's potential cell runoff of every continental cell of [p0, p1)
// (device order) to its running sum, G_potCellRunoff[n] += potCellRunoff
// (routing.cpp:1970, 2053, 2187, 3379): the local runoff terms, then the arid
// groundwater step's runoff where that step ran, then the open-water balance, in
// the reference's order of addition.
__global__ void wg_pot_runoff_kernel(WgStatic S, WgState X, int32_t p0, int32_t p1) {
  const int p = p0 + int(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= p1) return;
  const int f = S.cflags[p];
  if (!(f & CFL_CONT)) return;
  const int64_t mp = int64_t(blockIdx.y) * S.n + p;
  const double base = X.pot_base[mp];
  const double lake_balance = X.pot_ow[mp] * X.cfa[mp];  // (dailyPrec - dailyOpenWaterPET) * CFA
  const double ow = (lake_balance / 1000000.0) *
                    ((S.lake_area[p] + S.reservoir_area[p]) +
                     ((S.cell_area[p] / 100.0) * ((S.loc_lake[p] + S.loc_wetland[p]) + S.glo_wetland[p])));
  const double pot = (f & CFL_ARID) ? (base + X.pot_lgw[mp]) + ow : base + ow;
  X.pot_acc[mp] += pot;
}

// Sets the day clock read by the kernels of a captured block of days; launched
// before each replay so that one graph serves every block of the same length.
__global__ void wg_clock_set(WgClock *clock, int64_t day0, int32_t doy0, int32_t started) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    clock->day0 = day0;
    clock->doy0 = doy0;
    clock->started = started;
  }
}

// Evaluates the shared math functions on the device for a batch of arguments
// (0: wg_exp(x), 1: wg_pow(x, y), 2: sqrt(x), 3: wg_log(x)); used to show that
// the device returns the host's bits.
__global__ void wg_libm_kernel(int fn, int n, const double *__restrict__ x, const double *__restrict__ y,
                               double *__restrict__ out) {
  load_exp2_table();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r;
  switch (fn) {
    case 0: r = wg_exp(x[i]); break;
    case 1: r = wg_pow(x[i], y[i]); break;
    case 2: r = sqrt(x[i]); break;
    case 3: r = wg_log(x[i]); break;
    default: {
      double sn, cs;
      wg_sincos2pi(x[i], &sn, &cs);
      r = (fn == 4) ? sn : cs;
      break;
    }
  }
  out[i] = r;
}
