// wg_route_step.h -- one cell's routing step with upstream water (routing.cpp:
// 2623-3958), the calendar facts it reads, and the constants of the day loop.
//
// Plain C++ that the kernels compile for the device and tests/test_route_step.py
// compiles for the CPU, where it checks that the split step (step_pre + step_fast)
// gives chain_step's bits.  The includer provides fma and fabs, WgDev types from
// wg_device.h, and includes this file inside its own namespace.
#ifndef WG_ROUTE_STEP_H
#define WG_ROUTE_STEP_H

#if defined(__HIPCC__) || defined(__HIP__)
#define WG_RS_FN __device__ inline
#define WG_RS_CONST __constant__
#else
#define WG_RS_FN inline
#define WG_RS_CONST static const
#endif

constexpr double kCanopyEvapoExp = 0.66666666;  // daily.cpp:1867
constexpr double kRunoffFracBuiltUp = 0.5;      // daily.cpp:1868
constexpr double kOpenWaterAlbedo = 0.08;       // daily.cpp:1870
constexpr double kLakeOutflowExp = 1.5;         // routing.cpp:125
constexpr double kWetlOutflowExp = 2.5;         // routing.cpp:126
constexpr double kEvapoRedExp = 3.32193;        // routing.cpp:127
constexpr double kEvapoRedExpRes = 2.81383;     // routing.cpp:128
constexpr double kKswbgw = 10.;                 // routing.cpp:1662
constexpr double kMinStorVol = 1.e-15;          // routing.h:24

WG_RS_CONST short kFirstDayInMonth[12] = {1, 32, 60, 91, 121, 152, 182, 213, 244, 274, 305, 335};
WG_RS_CONST short kDaysInMonth[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};

// The calendar facts the routing step reads for one day of the year, computed once
// per launch from constant table entries (no table lookup with a run-time index
// remains in the per-cell code).
struct DayInfo {
  int month;              // 0-based
  bool first_of_month;    // the day is the month's first (routing.cpp:2956-2957)
  double days_in_month;
};

WG_RS_FN DayInfo day_info(int doy) {
  DayInfo d;
  d.month = 0;
  int first = kFirstDayInMonth[0], len = kDaysInMonth[0];
#pragma unroll
  for (int k = 1; k < 12; k++) {
    if (doy >= kFirstDayInMonth[k]) {
      d.month = k;
      first = kFirstDayInMonth[k];
      len = kDaysInMonth[k];
    }
  }
  d.first_of_month = (doy == first);
  d.days_in_month = double(len);
  return d;
}

// ---- the upstream-dependent part of one routing cell ----

// Record of one routing cell: every term of routing.cpp:2623-3958 that does not
// depend on the water arriving from upstream, computed by the cell kernel with
// the reference's own expressions, so that the level walk only evaluates the
// short chain that does.  Stored as one 304-byte record per (member, cell).
enum ChainField : int {
  // prefix read by every routed cell: river, global wetland, arid groundwater
  CF_INFLOW = 0,  // outflow of the local lake/wetland system (routing.cpp:2616)
  CF_LRIR,        // G_localRunoffIntoRiver
  CF_LGWRIR,      // G_localGWRunoffIntoRiver of humid cells
  CF_RV_PREC,     // precipitation on the river (routing.cpp:3457)
  CF_RV_TERM,     // K S0 e^-K / (1 - e^-K) (routing.cpp:3477-3478)
  CF_RV_EVAP,     // river evaporation (routing.cpp:3449-3453)
  CF_RV_S0E, CF_RV_INVK, CF_RV_OME, CF_RV_PREV,
  CF_GW_WP,       // precipitation on the global wetland (routing.cpp:3212-3213)
  CF_GW_PET, CF_GW_MAX, CF_GW_PREV, CF_GW_S0E, CF_GW_GWR,
  CF_A_GWR_LL, CF_A_GWR_LW, CF_A_RECH, CF_A_PREV, CF_A_S0E,
  CF_PREFIX,
  // suffix read only by cells with a global lake or a reservoir
  CF_GL_LP = CF_PREFIX,  // precipitation on the global lake (routing.cpp:2664)
  CF_GL_USE, CF_GL_MAX, CF_GL_PREV, CF_GL_S0E, CF_GL_GWR,
  CF_OW_PREC, CF_OW_PET,
  // reservoir inputs (routing.cpp:2807-2860): previous storage, evaporation
  // reduction factor, CFA, release coefficient, capacity, mean outflow, area, the
  // provisional release of an irrigation reservoir (routing.cpp:2972-2996 with
  // the mean demand of routing.cpp:360-403, computed once on the host), padding
  CF_RES_PREV, CF_RES_RED, CF_RES_CFA, CF_RES_KREL, CF_RES_CAP, CF_RES_QMEAN, CF_RES_AREA, CF_RES_PROVREL,
  CF_RES_PAD,
  // water use (loaded only when it is on; 0 otherwise): the day's desired
  // surface-water use (routing.cpp:2193-2290) and the net abstraction from the
  // groundwater of (semi-)arid cells (routing.cpp:3324-3330)
  CF_USE, CF_A_NUG,
  CF_COUNT
};
static_assert(CF_COUNT == WG_DEV_NREC, "routing record size");
static_assert(CF_COUNT % 2 == 0, "routing record stored in pairs");
static_assert(CF_USE % 2 == 0 && CF_A_NUG == CF_USE + 1 && CF_COUNT == CF_USE + 2, "water-use pair last");

// chain flags (static per cell); bits 8-15 reservoir type, 16-23 start month
enum {
  CFL_LAKE = 1, CFL_RES = 2, CFL_GWET = 4, CFL_ARID = 8, CFL_LDD = 16, CFL_CONT = 32
};
WG_RS_FN int flag_res_type(int flags) { return (flags >> 8) & 0xff; }
WG_RS_FN int flag_start_month(int flags) { return (flags >> 16) & 0xff; }

struct ChainRec {
  double v[CF_COUNT];
  double inv_ksw, ome_sw, inv_kgw, ome_gw, area, cf100;
  int flags;
};

// lgw: G_localGWRunoffIntoRiver after the step (the arid groundwater step's
// runoff in (semi-)arid cells), a term of the potential cell runoff
// rem: the surface-water use the river left (routing.cpp:3489-3512), which the
// local lake is asked for at the end of the cell's routing (3590-3622)
struct ChainOut {
  double glo_lake, res, k_release, glo_wetl, gw, river, tv, lgw, rem;
};

struct RiverOut {
  double st, tv, rem;
};

// River storage and transported volume, routing.cpp:3402-3514: riverInflow is the
// outflow of the cell's surface water bodies (or its inflow), plus local runoff
// and the precipitation on the river.
WG_RS_FN RiverOut river_step(double inflow, int flags, double lrir, double lgwrir, double rv_prec,
                                      double prev, double rv_evap, double rv_term, double s0e, double invk,
                                      double ome, double remainingUse) {
  double riverInflow = inflow;
  if (flags & CFL_LDD) {
    riverInflow += lrir;
    riverInflow += lgwrir;
  }
  riverInflow += rv_prec;
  const double evapRemUse = remainingUse + rv_evap;
  const double evapRemUseMax = riverInflow + rv_term;
  RiverOut r;
  if (evapRemUse > evapRemUseMax) {
    r.st = 0.;
    r.tv = riverInflow + prev - evapRemUseMax;
    if (r.tv < 0.) r.tv = 0.;
    if (remainingUse > 0.)
      r.rem = remainingUse - remainingUse * evapRemUseMax / evapRemUse;
    else
      r.rem = 0.;
  } else {
    r.st = s0e + invk * (riverInflow - evapRemUse) * ome;
    if (fabs(r.st) <= kMinStorVol) r.st = 0.;
    r.tv = riverInflow + prev - r.st - evapRemUse;
    if (r.tv < 0.) r.tv = 0.;
    r.rem = 0.;
  }
  return r;
}

// Groundwater of a (semi-)arid cell with recharge below surface water bodies,
// routing.cpp:3305-3386, less the day's net abstraction nug (0 without water use).
// Returns the groundwater runoff; gw receives the storage.
WG_RS_FN double arid_groundwater(double gwr_swb, double area, double cf100, double recharge,
                                         double gw_prev, double s0e, double inv_kgw, double ome_gw,
                                         double nug, double &gw_out) {
  double netGWin = gwr_swb * area * cf100 / 1000000. + recharge;
  netGWin -= nug;
  double gw = s0e + inv_kgw * netGWin * ome_gw;
  if (fabs(gw) <= kMinStorVol) gw = 0.;
  double gw_runoff = gw_prev - gw + netGWin;
  if (gw_runoff <= 0.) {
    gw_runoff = 0.;
    gw = gw_prev + netGWin;
    if (fabs(gw) <= kMinStorVol) gw = 0.;
  }
  gw_out = gw;
  return gw_runoff;
}

// routing.cpp:2623-3958 for one cell from its record (no station correction).
// inflow_up is G_riverInflow[n] when the cell is reached.  The reservoir branch
// (about one cell in two hundred) reads its own inputs here.  The water use
// (CF_USE) is shared between a global lake and a reservoir, taken from each
// (routing.cpp:2678-2716, 2866-2945), and what is left goes to the river; without
// water use it is +0 throughout, which leaves every result unchanged.
// kUse false: water use off, compiled with the use terms at +0.
template <bool kUse = true>
WG_RS_FN ChainOut chain_step(const ChainRec &c, double inflow_up, const DayInfo &d) {
  const int month = d.month;
  ChainOut o;
  const double *v = c.v;
  const bool arid_swb = (c.flags & CFL_ARID) != 0;
  double inflow = v[CF_INFLOW];
  double outflow = 0.;
  // the record suffix (CF_GL_*, CF_OW_*, CF_RES_*) is read only for cells with a
  // global lake or reservoir: other cells' suffix holds no data
  double gwr_glolak = (c.flags & CFL_LAKE) ? v[CF_GL_GWR] : 0., gwr_glowet = v[CF_GW_GWR], gwr_res = 0.;
  o.glo_lake = 0.;
  o.glo_wetl = v[CF_GW_PREV];
  o.gw = v[CF_A_PREV];
  o.res = 0.;
  o.k_release = 0.;

  inflow += inflow_up;  // routing.cpp:2623
  double remainingUse = kUse ? v[CF_USE] : 0., remainingUseGloLake = 0., remainingUseRes = 0.;

  // global lake, routing.cpp:2630-2804; CF_GL_USE holds the lake's share of the use
  if (c.flags & CFL_LAKE) {
    const double prev = v[CF_GL_PREV], use = v[CF_GL_USE], maxStorage = v[CF_GL_MAX];
    const double totalInflow = inflow + v[CF_GL_LP];
    const double useMax = totalInflow + maxStorage + prev;
    remainingUseGloLake = (c.flags & CFL_RES) ? 0.5 * remainingUse : remainingUse;
    double st;
    if (use > useMax) {
      st = (-1.) * maxStorage;
      outflow = 0.;
      gwr_glolak *= useMax / use;
      if (remainingUseGloLake > 0.)
        remainingUseGloLake -= remainingUseGloLake * useMax / use;
      else
        remainingUseGloLake = 0.;
    } else {
      st = v[CF_GL_S0E] + c.inv_ksw * (totalInflow - use) * c.ome_sw;
      outflow = totalInflow + prev - st - use;
      if (st > maxStorage) {
        outflow += (st - maxStorage);
        st = maxStorage;
      }
      if (outflow < 0.) {
        outflow = 0.;
        st = prev + totalInflow - use;
      }
      remainingUseGloLake = 0.;
    }
    if (fabs(st) <= kMinStorVol) st = 0.;
    o.glo_lake = st;
    inflow = outflow;
  }

  // reservoir, routing.cpp:2807-3082
  if (c.flags & CFL_RES) {
    const double res_area = v[CF_RES_AREA];
    const double prev = v[CF_RES_PREV], rf = v[CF_RES_RED], cfa = v[CF_RES_CFA];
    const double cap = v[CF_RES_CAP], qmean = v[CF_RES_QMEAN];
    const double ow_prec = v[CF_OW_PREC], ow_pet = v[CF_OW_PET];
    const double cellArea = c.area, cf100 = c.cf100;
    double evap, PETgwr, PETgwrMax, st, prov_rel = 0., release;
    const double c_ratio = cap / (qmean * 31536000. / 1000000000.);
    const double maxStorage = cap;
    evap = ((1.0 - cfa) * ow_prec) + (cfa * (ow_pet * rf));
    if (evap < 0.) evap = 0.;
    const double totalInflow = inflow + (ow_prec * (res_area / 1000000.));
    if (arid_swb)
      gwr_res = kKswbgw * rf * (res_area / (cellArea * cf100));
    else
      gwr_res = 0.;
    PETgwr = evap * (res_area / 1000000.) + gwr_res * cellArea * cf100 / 1000000.;
    PETgwrMax = prev + totalInflow;
    if (PETgwr > PETgwrMax) {
      st = prev + totalInflow - PETgwrMax;
      gwr_res *= PETgwrMax / PETgwr;
    } else {
      st = prev + totalInflow - PETgwr;
    }
    // the use down to 10 % of the capacity, a return flow (< 0) added
    // (routing.cpp:2869-2874, 2922-2945); with +0 use the storage is unchanged
    remainingUseRes = (c.flags & CFL_LAKE) ? 0.5 * remainingUse + remainingUseGloLake : remainingUse;
    if (remainingUseRes < 0.) {
      st -= remainingUseRes;
      remainingUseRes = 0.;
    } else if (st > (0.1 * cap)) {
      if (remainingUseRes < (st - (cap * 0.1))) {
        st -= remainingUseRes;
        remainingUseRes = 0.;
      } else {
        remainingUseRes -= (st - (cap * 0.1));
        st = cap * 0.1;
      }
    }
    if (fabs(st) <= kMinStorVol) st = 0.;
    double krel = v[CF_RES_KREL];
    if (month == flag_start_month(c.flags) - 1) {
      if (d.first_of_month) {
        if (st < (cap * 0.1))
          krel = 0.1;
        else
          krel = st / (maxStorage * 0.85);
      }
    }
    const int rtype = flag_res_type(c.flags);
    if (rtype == 1) {
      // irrigation reservoir: the month's provisional release (routing.cpp:2970-2996)
      prov_rel = v[CF_RES_PROVREL];
    } else if (rtype == 2) {
      prov_rel = qmean;
    }
    if (c_ratio >= 0.5)
      release = krel * prov_rel;
    else
      release = ((4. * c_ratio * c_ratio) * krel * prov_rel) +
                ((1.0 - ((4. * c_ratio * c_ratio))) * inflow * 1000000000. / (24. * 3600.));
    if (st >= (cap * 0.1))
      outflow = release * (24. * 3600.) / 1000000000.;
    else
      outflow = 0.1 * release * (24. * 3600.) / 1000000000.;
    if (outflow < 0.) outflow = 0.;
    st -= outflow;
    if (st > maxStorage) {
      outflow += (st - maxStorage);
      st = maxStorage;
    }
    if (st < 0.) {
      outflow += st;
      st = 0.;
    }
    o.res = st;
    o.k_release = krel;
    inflow = outflow;
  }

  // routing.cpp:3162-3171 (aggrNUsGloLakResOpt = 0)
  if (c.flags & CFL_RES)
    remainingUse = remainingUseRes;
  else if (c.flags & CFL_LAKE)
    remainingUse = remainingUseGloLake;

  // global wetland, routing.cpp:3178-3297
  if (c.flags & CFL_GWET) {
    const double prev = v[CF_GW_PREV], PETgwr = v[CF_GW_PET], maxStorage = v[CF_GW_MAX];
    const double totalInflow = inflow + v[CF_GW_WP];
    const double PETgwrMax = totalInflow + prev;
    double st;
    if (PETgwr > PETgwrMax) {
      st = 0.;
      outflow = 0.;
      gwr_glowet *= PETgwrMax / PETgwr;
    } else {
      st = v[CF_GW_S0E] + c.inv_ksw * (totalInflow - PETgwr) * c.ome_sw;
      outflow = totalInflow + prev - st - PETgwr;
    }
    if (st > maxStorage) {
      outflow += (st - maxStorage);
      st = maxStorage;
    }
    if (fabs(st) <= kMinStorVol) st = 0.;
    o.glo_wetl = st;
    inflow = outflow;
  }

  // groundwater of (semi-)arid cells, routing.cpp:3305-3386
  double lgwrir = v[CF_LGWRIR];
  if (arid_swb) {
    const double gwr_swb = v[CF_A_GWR_LL] + gwr_glolak + v[CF_A_GWR_LW] + gwr_glowet + gwr_res;
    lgwrir = arid_groundwater(gwr_swb, c.area, c.cf100, v[CF_A_RECH], v[CF_A_PREV], v[CF_A_S0E], c.inv_kgw,
                              c.ome_gw, kUse ? v[CF_A_NUG] : 0., o.gw);
  }

  // river, routing.cpp:3402-3514
  const RiverOut r = river_step(inflow, c.flags, v[CF_LRIR], lgwrir, v[CF_RV_PREC], v[CF_RV_PREV], v[CF_RV_EVAP],
                                v[CF_RV_TERM], v[CF_RV_S0E], v[CF_RV_INVK], v[CF_RV_OME], remainingUse);
  const double st = r.st, tv = r.tv;
  o.river = st;
  o.tv = tv;
  o.lgw = lgwrir;
  o.rem = r.rem;
  return o;
}

// The common case of the routing step, split so that what does not depend on the
// upstream inflow is done beforehand: a cell without a reservoir whose global lake
// and global wetland (if any) keep water.  In that case the arid groundwater step
// (routing.cpp:3305-3386) does not depend on the inflow (step_pre), and the
// inflow-dependent rest (step_fast) is the lake, wetland and river steps, the
// latter two with selects in place of branches.  Every operation is the one
// chain_step performs, in the same order, so the results are chain_step's bits.
// step_fast flags the cells that need chain_step: a reservoir, or a lake or
// wetland that runs dry (which rescales its groundwater recharge,
// routing.cpp:2708 and 3235).
struct StepPre {
  double lgwrir, gw;
};

template <bool kUse = true>
WG_RS_FN StepPre step_pre(const ChainRec &c) {
  const double *v = c.v;
  StepPre r{v[CF_LGWRIR], v[CF_A_PREV]};
  if (c.flags & CFL_ARID) {
    // chain_step's sum with the lake's and wetland's recharge unscaled and no reservoir
    const double gwr_glolak = (c.flags & CFL_LAKE) ? v[CF_GL_GWR] : 0.;
    const double gwr_swb = v[CF_A_GWR_LL] + gwr_glolak + v[CF_A_GWR_LW] + v[CF_GW_GWR] + 0.;
    r.lgwrir = arid_groundwater(gwr_swb, c.area, c.cf100, v[CF_A_RECH], v[CF_A_PREV], v[CF_A_S0E], c.inv_kgw,
                                c.ome_gw, kUse ? v[CF_A_NUG] : 0., r.gw);
  }
  return r;
}

template <bool kUse = true>
WG_RS_FN ChainOut step_fast(const ChainRec &c, const StepPre &pre, double inflow_up, bool &slow) {
  const double *v = c.v;
  ChainOut o;
  o.glo_lake = 0.;
  o.res = 0.;
  o.k_release = 0.;
  o.gw = pre.gw;
  o.lgw = pre.lgwrir;
  double inflow = v[CF_INFLOW];
  inflow += inflow_up;  // routing.cpp:2623
  slow = (c.flags & CFL_RES) != 0;
  if (c.flags & CFL_LAKE) {
    // global lake, routing.cpp:2630-2804, the branch that keeps water
    const double prev = v[CF_GL_PREV], use = v[CF_GL_USE], maxStorage = v[CF_GL_MAX];
    const double totalInflow = inflow + v[CF_GL_LP];
    const double useMax = totalInflow + maxStorage + prev;
    slow = slow || (use > useMax);
    double st = v[CF_GL_S0E] + c.inv_ksw * (totalInflow - use) * c.ome_sw;
    double outflow = totalInflow + prev - st - use;
    const bool over = st > maxStorage;
    outflow = over ? outflow + (st - maxStorage) : outflow;
    st = over ? maxStorage : st;
    const bool neg = outflow < 0.;
    st = neg ? prev + totalInflow - use : st;
    outflow = neg ? 0. : outflow;
    st = (fabs(st) <= kMinStorVol) ? 0. : st;
    o.glo_lake = st;
    inflow = outflow;
  }
  {
    // global wetland, routing.cpp:3178-3297, the branch that keeps water
    const double prev = v[CF_GW_PREV], PETgwr = v[CF_GW_PET], maxStorage = v[CF_GW_MAX];
    const double totalInflow = inflow + v[CF_GW_WP];
    const double PETgwrMax = totalInflow + prev;
    double st = v[CF_GW_S0E] + c.inv_ksw * (totalInflow - PETgwr) * c.ome_sw;
    double outflow = totalInflow + prev - st - PETgwr;
    const bool over = st > maxStorage;
    outflow = over ? outflow + (st - maxStorage) : outflow;
    st = over ? maxStorage : st;
    st = (fabs(st) <= kMinStorVol) ? 0. : st;
    const bool gwet = (c.flags & CFL_GWET) != 0;
    slow = slow || (gwet && PETgwr > PETgwrMax);
    o.glo_wetl = gwet ? st : prev;
    inflow = gwet ? outflow : inflow;
  }
  {
    // river, routing.cpp:3402-3514 (river_step)
    const double prev = v[CF_RV_PREV];
    double rin = inflow;
    const double rin_ldd = (rin + v[CF_LRIR]) + pre.lgwrir;
    rin = (c.flags & CFL_LDD) ? rin_ldd : rin;
    rin += v[CF_RV_PREC];
    // a global lake that keeps water leaves no use (routing.cpp:2752); no reservoir here
    const double remainingUse = (kUse && !(c.flags & CFL_LAKE)) ? v[CF_USE] : 0.;
    const double evapRemUse = remainingUse + v[CF_RV_EVAP];
    const double evapRemUseMax = rin + v[CF_RV_TERM];
    double tv_dry = rin + prev - evapRemUseMax;
    tv_dry = (tv_dry < 0.) ? 0. : tv_dry;
    double st = v[CF_RV_S0E] + v[CF_RV_INVK] * (rin - evapRemUse) * v[CF_RV_OME];
    st = (fabs(st) <= kMinStorVol) ? 0. : st;
    double tv = rin + prev - st - evapRemUse;
    tv = (tv < 0.) ? 0. : tv;
    const bool dry = evapRemUse > evapRemUseMax;
    o.river = dry ? 0. : st;
    o.tv = dry ? tv_dry : tv;
    const double rem_dry = (remainingUse > 0.) ? remainingUse - remainingUse * evapRemUseMax / evapRemUse : 0.;
    o.rem = dry ? rem_dry : 0.;
  }
  return o;
}

#endif  // WG_ROUTE_STEP_H
