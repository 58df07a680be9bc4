// wg_device.h -- device data layout shared by the kernels and the host API.
//
// Cells live in a device order p (routing group, then Kahn level, then routing
// order), so that every level of a group is a contiguous range.  Per-member
// arrays are [member][N]; sub-cell arrays are [member][sub-cell pair][N] so that the
// 64 lanes of a wavefront (64 consecutive cells) read two sub-cells contiguously.
#ifndef WG_DEVICE_H
#define WG_DEVICE_H

#include <stdint.h>

#define WG_DEV_NSUB 100
#define WG_DEV_NPAIR 50

// Two consecutive elevation sub-cells of one cell (one 16-byte load).
struct alignas(16) WgD2 {
  double x, y;
};

// Routing links of a cell (device order): chain flags in bits 0-7, reservoir type
// in bits 8-15 and the reservoir's start month in bits 16-23; the station slot;
// the number of upstream cells and the largest such number on the cell's level;
// the cell's Kahn level counted from its group's first level; the upstream cells'
// group-local indices in routing order (at most 8, D8).
struct alignas(16) WgChainLink {
  int32_t flags, station;
  int16_t nu, nu_level;
  int32_t level;
  int16_t up[8];
};

// A batch of the wave walk's wide levels: cells [start, end) of one level (counted
// from the group's first level), up to 64; lakeres is 1 when one of them has a
// global lake or a reservoir (the batch then needs the routing record's suffix).
struct alignas(16) WgWaveItem {
  int32_t start, end, level, lakeres;
};

// Outflow constants of the storages on the routing chain (global lake / wetland,
// arid groundwater).
struct alignas(16) WgChainStatic {
  double inv_ksw, ome_sw, inv_kgw, ome_gw, area, cf100;
};

// Routing record of a cell with upstream inflow, written by the cell kernel: a
// prefix read by every routed cell and a suffix written and read only for cells
// with a global lake or a reservoir (field meanings: ChainField in wg_kernels.hip).
// Stored as pairs of fields, [M][WG_DEV_NREC / 2][N] x 16 bytes, so that the lanes
// of a wavefront, which hold consecutive cells, read and write whole cache lines
// with 16-byte accesses.
#define WG_DEV_NREC 40

// Device layout of the routing links, [2][N] x 16 bytes: (flags, station,
// nu | nu_level << 16, level), then four words of two upstream indices (up[2k] in
// the low half); of the chain constants, [3][N] x 16 bytes in the order of
// WgChainStatic.
struct alignas(16) WgI4 {
  int32_t x, y, z, w;
};

// Land-cover tables indexed by class-1 (LCT_22.DAT, LAI_22.DAT).
struct WgTables {
  double albedo[18], albedo_snow[18], ddf[18], emissivity[18];
  float lai_a[18], lai_b[18];
  int32_t initial_days[18];
};

// Static per-cell fields (device order).  Derived geometry is computed on the
// host exactly as routingClass::init does.
struct WgStatic {
  int32_t n;            // cells
  int32_t n_members;
  const int32_t *ref_index;     // reference cell index of device cell p
  const double *cell_area, *contfreq;
  const int8_t *ldd, *land_cover, *texture, *contcell;
  const int16_t *aindex, *rgmax;
  const float *built_up, *gw_factor, *lai_max;
  const int16_t *elev0;         // [N] mean cell elevation (channel 0 of G_Elevation)
  const int32_t *elev_pairs;    // [50][N] sub-cell elevations 2k+1 (low half) and 2k+2 (high half)
  // calibration parameters used by the day loop
  const double *alpha, *pet_mxdy, *m_prec, *m_netrad, *mcwh, *t_snowfz, *t_snowmt, *t_gradnt,
      *degday, *pcritgwa, *evarexp, *m_rivrgh;
  // routing statics
  const double *loc_lake, *loc_wetland, *glo_wetland, *lake_area, *reservoir_area, *stor_cap,
      *mean_outflow, *res_prov_rel, *river_length, *river_slope, *roughness, *width_bf, *bottom_width,
      *river_stor_max, *lake_depth, *wetl_depth, *f_glo_lake, *fswb_init;
  const int8_t *res_type, *start_month;
  // exp(-k) terms of the surface-water and groundwater outflow coefficients
  const double *k_sw, *e_sw, *inv_ksw, *ome_sw, *k_gw, *e_gw, *inv_kgw, *ome_gw;
  // per-cell constants of the river velocity and width (routing.cpp:3546-3560, 7309-7341):
  // 1/(M_RIVRGH_C roughness), slope^0.5, -B/4000, (B/1000) B/16000, B/1000, width_bf/1000
  const double *vel_c1, *vel_c2, *rw_c0, *rw_c1, *rw_c2, *rw_c3;
  const double *cf100;              // contfreq / 100
  const int8_t *cflags;             // lake / reservoir / global wetland / arid / LDD >= 0 / continental
  const WgI4 *clink;                // [2][N] routing links
  const WgD2 *cstat;                // [3][N] outflow constants of chain storages
  // XCD-aware placement: the cell kernel's block b covers cells [tile_start[b],
  // tile_end[b]) (tile_start -1: idle block); the chain kernel's block i routes
  // group block_group[i] (-1: idle).  Block i runs on XCD i % 8.
  const int32_t *tile_start, *tile_end, *block_group;
  // schedule
  const int32_t *up_off, *up_idx;   // CSR of upstream cells sorted by routing order
  const int32_t *station;           // station slot of each cell, -1 = none
  int32_t n_stations;
  // water use (use_on 0: off).  Daily values by model year and month,
  // [use_years][12][N]: net use from surface water and from groundwater, the
  // irrigation withdrawal and consumptive use from surface water, and the
  // provisional release of irrigation reservoirs (m3/s); G_fractreturngw_irrig [N].
  // Model year y covers the forcing days use_day0 + 365 y .. + 364.
  int32_t use_on, use_delayed, use_years;
  int64_t use_day0;
  int64_t use_stride;           // M * N, between the fields of WgState::use
  const double *use_nus, *use_nug, *use_wu, *use_cu, *use_provrel, *use_frgi;
};

// Per-member prognostic state and day buffers ([member][N] unless noted).
struct WgState {
  double *canopy, *soil, *snow, *transfer, *lai_precsum;
  int32_t *lai_days, *lai_status;
  WgD2 *snow2;                  // [M][50][N] sub-cell snow storages 2k+1, 2k+2
  uint8_t *snow_any;            // [M][N] 0: every sub-cell storage is exactly 0
  double *gw, *loc_lake, *loc_wetl, *glo_lake, *glo_wetl, *res, *river;
  double *red_ll, *red_lw, *red_gl, *red_gw, *red_res, *red_river;
  double *river_frac_next, *river_frac_change, *fswb_next, *laf, *laf_prev, *k_release;
  double *gamma, *cfa;
  float *smax;
  double *tv;                   // transported volume of the day
  // routing record written by the cell kernel for cells with upstream inflow
  WgD2 *crec;                   // [M][WG_DEV_NREC / 2][N] routing records
  // diagnostics of the last day
  double *d_surf, *d_recharge, *d_ow_pet, *d_aet, *d_prec;
  // potential cell runoff (routing.cpp:1962-1970, 2045-2053, 2180-2187, 3372-3379),
  // allocated only when that output is on: the running sum since the last reset,
  // and the day's terms of each cell that the chain kernels complete
  double *pot_acc;              // [M][N] sum of the daily potential cell runoff (km3)
  double *pot_base;             // [M][N] local runoff terms of the day
  double *pot_ow;               // [M][N] open-water balance term of the day
  double *pot_lgw;              // [M][N] arid groundwater runoff of the day
  // water-use state [WG_USE_NFIELDS + 1][M][N] (field order of the C ABI's
  // WG_USE_*), the last field the use the river left on the day just routed,
  // which the next cell-kernel pass takes from the local lake
  double *use;
  double *use_rem;              // the last field of use (NULL: water use off)
  // stations
  double *st_sum;               // [M][n_stations]
  double *st_daily;             // [M][WG_DAILY_SLOTS][n_stations]
  int32_t *flow_error;           // [1] set when the dataflow chain walk hits its spin limit
  // forcing perturbation
  const uint64_t *member_seed;
  double prec_sigma, temp_sigma;
  int32_t perturb;
};

#define WG_DAILY_SLOTS 366
#define WG_MAX_GROUP_DEPTH 4096   // Kahn levels of one routing group (dataflow chain kernel)

#endif
