// wg_api.hip -- host side of the C ABI (include/watergap2_amd.h).
//
// wg_create validates the inputs, derives the routing geometry and the initial
// state the way the reference initialises a run (routingClass::init
// routing.cpp:131-742, initFractionStatus 745-765, setStoragesToZero 789-818,
// setLakeWetlToMaximum 5647-5720, annualInit 979-1495, laiClass::init
// lai.cpp:40-108), builds the level schedule, permutes every array into the
// device order and uploads it.  wg_step_days replays hipGraphs that hold one
// block of days each (two kernels per day plus a closing pass).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/watergap2_amd.h"
#include "wg_device.h"
// The kernels are compiled in this translation unit so that their launches need
// no relocatable device code.  (wg_kernels.hip includes wg_libm.h.)
#include "wg_kernels.hip"

namespace {


struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

// One partition of the routing groups.  With two partitions the deepest groups
// form the first; each partition's days run as a graph of its own on a stream of
// its own, so that the long walk of the deep groups overlaps the other groups'
// days (basins are independent, so the results do not change).  A partition's
// cells, tiles and chain blocks are contiguous ranges of the whole grid's.
struct WgPart {
  std::vector<int32_t> bin_start, bin_end;  // device cell range of each XCD bin
  int32_t cell_begin = 0, cell_end = 0;     // device cell range
  int32_t tile_off = 0, n_tiles = 0;        // range of the whole grid's tile table
  int32_t group_block_off = 0, n_group_blocks = 0;  // range of its chain block table
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::map<int, hipGraphExec_t> graphs;     // keyed by block length
};

}  // namespace

struct wg_handle {
  int32_t n = 0, m = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<DevBuf> bufs;
  // schedule
  std::vector<int32_t> perm;      // device index -> reference index
  std::vector<int32_t> inv;       // reference index -> device index
  std::vector<int32_t> level_off; // global level slots -> start p (plus sentinel)
  std::vector<int32_t> group_first_level;
  int32_t n_groups = 0, max_depth = 0;
  int32_t *d_level_off = nullptr, *d_group_first = nullptr;
  // wave walk: the number of wide levels (more than 64 cells) leading each group
  // above level 0, and their batches dealt to W = 4 or 8 wavefronts (tables [0]
  // and [1]): the batches of wavefront w of group g are
  // d_wave_items[i][d_wave_item_off[i][W g + w] .. [W g + w + 1])
  int32_t *d_wave_wide = nullptr, *d_wave_item_off[2] = {nullptr, nullptr};
  WgWaveItem *d_wave_items[2] = {nullptr, nullptr};
  // host copies needed at the boundary
  std::vector<double> snow_ch0;   // [M][N] channel 0 of G_SnowInElevation (never read by calcNewDay)
  WgStatic S{};
  WgState X{};
  WgTables T{};
  WgClock *d_clock = nullptr;
  bool started = false;
  // forcing
  float *d_forcing = nullptr;
  int64_t n_forcing_days = 0;
  // stations
  int32_t n_stations = 0;
  std::vector<int32_t> station_cells;
  int32_t last_daily_days = 0;   // days of the daily station series kept from the last step call
  int32_t last_daily_first = 0;  // ... and the first of them, counted from that call's first day
  // graphs keyed by block length
  std::map<int, hipGraphExec_t> graphs;
  int block_threads_cell = 256, block_threads_chain = 256;
  int snow_chunk = 5;             // sub-cell pairs per double-buffered chunk of the cell kernel
  int chain_mode = 3;             // chain kernel: 0 level walk, loads one level ahead; 1 dataflow walk;
                                  // 2 level walk, loads two levels ahead; 3 wave walk in batches
  int32_t max_group_cells = 0;
  int n_xcd = 8;
  std::vector<WgPart> parts;                 // at least one; two run concurrently
  // water use: the year tables' generation (wg_set_water_use), which must follow
  // the last wg_annual_init (the irrigation reservoirs' release depends on both)
  int32_t statics_gen = 0, use_gen = -1;
  int32_t subtract_use = 0;
  hipEvent_t ev_fork = nullptr;
  std::vector<int32_t> tile_start, tile_end, block_group;  // whole grid: the partitions' tables in turn
  int32_t tile_capacity = 0;      // entries of the device tile tables
  std::vector<WgChainLink> clink;
  std::vector<int32_t> wave_wide;    // wide levels of each group (host copy of d_wave_wide)
  std::vector<double> glo_res_prev;  // G_glores_prevyear, reference order (routing.cpp:5479-5483)
  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms[3] = {0, 0, 0};
  int64_t last_launches[3] = {0, 0, 0};
  uint64_t *d_seeds = nullptr;
};

namespace {

int fail(wg_handle *h, int code, const char *fmt, ...) {
  if (h) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    h->err = buf;
  }
  return code;
}

#define HIP_CHECK(h, call)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((h), WG_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_));            \
  } while (0)

template <class T>
int dev_alloc(wg_handle *h, T **out, size_t count) {
  void *p = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return fail(h, WG_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  // The handle's stream does not synchronise with the null stream that the
  // uploads use: finish the zero fill before any copy can target the buffer.
  e = hipMemsetAsync(p, 0, bytes, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    (void)hipFree(p);
    return fail(h, WG_E_HIP, "zero fill failed: %s", hipGetErrorString(e));
  }
  h->bufs.push_back({p, bytes});
  *out = static_cast<T *>(p);
  return WG_OK;
}

// Frees one buffer that dev_alloc allocated.
void release_buf(wg_handle *h, void *ptr) {
  for (auto it = h->bufs.begin(); it != h->bufs.end(); ++it)
    if (it->p == ptr) {
      (void)hipFree(it->p);
      h->bufs.erase(it);
      return;
    }
}

template <class T>
int upload(wg_handle *h, T **out, const std::vector<T> &host) {
  int rc = dev_alloc(h, out, host.size());
  if (rc) return rc;
  if (!host.empty()) {
    hipError_t e = hipMemcpy(*out, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(h, WG_E_HIP, "hipMemcpy failed: %s", hipGetErrorString(e));
  }
  return WG_OK;
}

// Routing links in the device layout ([2][N] x 16 bytes).
std::vector<WgI4> pack_links(const std::vector<WgChainLink> &lk) {
  const size_t n = lk.size();
  std::vector<WgI4> w(2 * n);
  auto pair16 = [](int16_t lo, int16_t hi) { return int32_t(uint32_t(uint16_t(lo)) | (uint32_t(uint16_t(hi)) << 16)); };
  for (size_t p = 0; p < n; p++) {
    const WgChainLink &l = lk[p];
    w[p] = WgI4{l.flags, l.station, pair16(l.nu, l.nu_level), l.level};
    w[n + p] = WgI4{pair16(l.up[0], l.up[1]), pair16(l.up[2], l.up[3]), pair16(l.up[4], l.up[5]),
                    pair16(l.up[6], l.up[7])};
  }
  return w;
}

// host array in reference order -> device order
template <class T>
std::vector<T> permute(const wg_handle *h, const T *ref) {
  std::vector<T> out(h->n);
  for (int32_t p = 0; p < h->n; p++) out[p] = ref[h->perm[p]];
  return out;
}

// per-member replication of a device-order vector
template <class T>
std::vector<T> replicate(const std::vector<T> &v, int m) {
  std::vector<T> out;
  out.reserve(v.size() * m);
  for (int k = 0; k < m; k++) out.insert(out.end(), v.begin(), v.end());
  return out;
}

bool options_supported(const wg_options_profile *o) {
  return o->time_series == 1 && o->cloud == 1 && o->intercept == 1 && o->calc_albedo == 0 &&
         o->pet_opt == 0 && o->use_kc == 0 && o->river_velo_opt == 2 &&
         (o->subtract_use == 0 || ((o->subtract_use == 2 || o->subtract_use == 3) && o->aggr_nus_glolakres_opt == 0)) &&
         o->res_opt == 1 && o->statcorr_opt == 0 && o->arid_area_opt == 1 && o->fractional_routing_opt == 1 &&
         o->river_evapo_opt == 1 && (o->aggr_nus_glolakres_opt == 0 || o->aggr_nus_glolakres_opt == 1);
}

// The per-member state arrays in the order of the wg_state struct.
struct StateMap {
  const char *name;
  double *wg_state::*ref_d;
  int32_t *wg_state::*ref_i;
  double *WgState::*dev_d;
  int32_t *WgState::*dev_i;
};

const StateMap kStateMap[] = {
    {"canopy", &wg_state::canopy, nullptr, &WgState::canopy, nullptr},
    {"soil", &wg_state::soil, nullptr, &WgState::soil, nullptr},
    {"snow", &wg_state::snow, nullptr, &WgState::snow, nullptr},
    {"storage_transfer", &wg_state::storage_transfer, nullptr, &WgState::transfer, nullptr},
    {"lai_days", nullptr, &wg_state::lai_days, nullptr, &WgState::lai_days},
    {"lai_status", nullptr, &wg_state::lai_status, nullptr, &WgState::lai_status},
    {"lai_precsum", &wg_state::lai_precsum, nullptr, &WgState::lai_precsum, nullptr},
    {"groundwater", &wg_state::groundwater, nullptr, &WgState::gw, nullptr},
    {"loc_lake", &wg_state::loc_lake, nullptr, &WgState::loc_lake, nullptr},
    {"loc_wetl", &wg_state::loc_wetl, nullptr, &WgState::loc_wetl, nullptr},
    {"glo_lake", &wg_state::glo_lake, nullptr, &WgState::glo_lake, nullptr},
    {"glo_wetl", &wg_state::glo_wetl, nullptr, &WgState::glo_wetl, nullptr},
    {"reservoir", &wg_state::reservoir, nullptr, &WgState::res, nullptr},
    {"river", &wg_state::river, nullptr, &WgState::river, nullptr},
    {"red_loc_lake", &wg_state::red_loc_lake, nullptr, &WgState::red_ll, nullptr},
    {"red_loc_wetl", &wg_state::red_loc_wetl, nullptr, &WgState::red_lw, nullptr},
    {"red_glo_lake", &wg_state::red_glo_lake, nullptr, &WgState::red_gl, nullptr},
    {"red_glo_wetl", &wg_state::red_glo_wetl, nullptr, &WgState::red_gw, nullptr},
    {"red_res", &wg_state::red_res, nullptr, &WgState::red_res, nullptr},
    {"red_river", &wg_state::red_river, nullptr, &WgState::red_river, nullptr},
    {"river_frac_next", &wg_state::river_frac_next, nullptr, &WgState::river_frac_next, nullptr},
    {"fswb_frac_next", &wg_state::fswb_frac_next, nullptr, &WgState::fswb_next, nullptr},
    {"laf", &wg_state::laf, nullptr, &WgState::laf, nullptr},
    {"laf_prev", &wg_state::laf_prev, nullptr, &WgState::laf_prev, nullptr},
    {"k_release", &wg_state::k_release, nullptr, &WgState::k_release, nullptr},
};

}  // namespace

extern "C" void wg_default_options(wg_options_profile *o) {
  if (!o) return;
  o->time_series = 1;
  o->cloud = 1;
  o->intercept = 1;
  o->calc_albedo = 0;
  o->pet_opt = 0;
  o->use_kc = 0;
  o->river_velo_opt = 2;
  o->subtract_use = 0;
  o->res_opt = 1;
  o->statcorr_opt = 0;
  o->arid_area_opt = 1;
  o->fractional_routing_opt = 1;
  o->river_evapo_opt = 1;
  o->aggr_nus_glolakres_opt = 0;
}

extern "C" const char *wg_last_error(const wg_handle *h) { return h ? h->err.c_str() : "null handle"; }

namespace {

// LDS of the wave walk: the transported volume (8 B) and the chain flags (1 B) of
// every cell of the largest routing group.
size_t wave_lds_bytes(const wg_handle *h) { return (size_t(h->max_group_cells) * 9 + 15) / 16 * 16; }

// Provisional release (m3/s) of each reservoir cell for the release rule of
// routing.cpp:2972-2999 in a month of `days` days whose daily net surface-water
// use (km3/day, reference cell order) is daily_nus; without water use (NULL) the
// use of the cell and of its downstream cells is 0, the month's use term +0 and
// the value the same every day:
//  - irrigation reservoirs (type 1): from the mean demand of the cell and of up to
//    five downstream cells, routing.cpp:360-403: the cell's mean net surface-water
//    use plus each downstream cell's weighted by its allocation coefficient,
//    walking down until a cell with a reservoir, in m3/yr / 31536000 = m3/s.  The
//    reference tests the reservoir area at G_reservoir_area_full[downstreamCell]
//    with the 1-based downstream cell number, i.e. the area of the cell after the
//    downstream cell (routing.cpp:372); the restatement keeps that and treats an
//    index past the last cell as a cell without a reservoir.  Without net-use
//    input the demand is 0.  Allocation coefficients default to 1, the value
//    rout_prepare.cpp:906-909 gives cells that no reservoir allocates to.
//  - other reservoirs (type 2): the mean outflow.
std::vector<double> reservoir_prov_rel(const wg_static_fields *st, int32_t n, const double *daily_nus = nullptr,
                                       int days = 31) {
  std::vector<double> out(n, 0.);
  for (int32_t i = 0; i < n; i++) {
    if (!(st->reservoir_area[i] > 0.)) continue;
    const double qmean = st->mean_outflow[i];
    if (st->res_type[i] == 2) {
      out[i] = qmean;
      continue;
    }
    double mean_demand = 0.;
    if (st->mean_nus) {
      mean_demand += st->mean_nus[i];
      int k = 0;
      int32_t dc = st->downstream[i];
      while (k < 5 && dc > 0 && (dc < n ? st->reservoir_area[dc] : 0.) <= 0) {
        mean_demand += st->mean_nus[dc - 1] * (st->alloc_coeff ? st->alloc_coeff[size_t(i) * 5 + k] : 1.0);
        k++;
        dc = st->downstream[dc - 1];
      }
    }
    mean_demand = mean_demand / 31536000.;
    // routing.cpp:2972-2996: the cell's daily use and that of up to five
    // downstream cells, walked as for the mean demand
    double daily_use = 0.0;
    if (daily_nus) {
      daily_use = daily_nus[i];
      int k = 0;
      int32_t dc = st->downstream[i];
      while (k < 5 && dc > 0 && (dc < n ? st->reservoir_area[dc] : 0.) <= 0) {
        daily_use += daily_nus[dc - 1] * (st->alloc_coeff ? st->alloc_coeff[size_t(i) * 5 + k] : 1.0);
        k++;
        dc = st->downstream[dc - 1];
      }
    }
    double monthly_use = daily_use * double(days);
    monthly_use = monthly_use * 1000000000. / (double(days) * 86400.);
    if (mean_demand >= 0.5 * qmean)
      out[i] = qmean / 2. * (1. + monthly_use / mean_demand);
    else
      out[i] = qmean + monthly_use - mean_demand;
  }
  return out;
}

// Cell-kernel tiles of block_threads_cell cells: within each partition, block
// 8j + x covers tile j of XCD bin x; the whole grid's table holds the partitions'
// tables in turn (each a multiple of 8 blocks long, so placement is kept).
void build_tiles(wg_handle *h) {
  const int nx = h->n_xcd, T = h->block_threads_cell;
  h->tile_start.clear();
  h->tile_end.clear();
  for (WgPart &P : h->parts) {
    int32_t max_tiles = 1;
    for (int x = 0; x < nx; x++) max_tiles = std::max(max_tiles, (P.bin_end[x] - P.bin_start[x] + T - 1) / T);
    P.tile_off = int32_t(h->tile_start.size());
    P.n_tiles = nx * max_tiles;
    h->tile_start.resize(size_t(P.tile_off) + P.n_tiles, -1);
    h->tile_end.resize(size_t(P.tile_off) + P.n_tiles, -1);
    for (int32_t j = 0; j < max_tiles; j++)
      for (int x = 0; x < nx; x++) {
        const int32_t a = P.bin_start[x] + T * j;
        if (a < P.bin_end[x]) {
          h->tile_start[size_t(P.tile_off) + size_t(j) * nx + x] = a;
          h->tile_end[size_t(P.tile_off) + size_t(j) * nx + x] = std::min(a + T, P.bin_end[x]);
        }
      }
  }
}

// Builds the level schedule and the device order.
int build_schedule(wg_handle *h, const wg_static_fields *st, int group_cells) {
  const int32_t n = h->n;
  std::vector<int32_t> by_order(n, -1);
  for (int32_t i = 0; i < n; i++) {
    const int32_t o = st->rout_order[i];
    if (o < 1 || o > n || by_order[o - 1] >= 0)
      return fail(h, WG_E_INPUT, "rout_order is not a permutation of 1..N (cell %d)", i);
    by_order[o - 1] = i;
  }
  for (int32_t i = 0; i < n; i++) {
    const int32_t d = st->downstream[i];
    if (d < 0 || d > n) return fail(h, WG_E_INPUT, "downstream cell of %d out of range", i);
    if (d > 0 && st->rout_order[d - 1] <= st->rout_order[i])
      return fail(h, WG_E_CYCLE, "routing order is not upstream-to-downstream at cell %d", i);
  }
  // Kahn level: 0 without upstream cells, else 1 + max level upstream
  std::vector<int32_t> level(n, 0), outlet(n, -1);
  for (int32_t k = 0; k < n; k++) {
    const int32_t i = by_order[k], d = st->downstream[i];
    if (d > 0) level[d - 1] = std::max(level[d - 1], level[i] + 1);
  }
  for (int32_t k = n - 1; k >= 0; k--) {
    const int32_t i = by_order[k], d = st->downstream[i];
    outlet[i] = d > 0 ? outlet[d - 1] : i;
  }
  // basins: outlet -> (depth, cells)
  std::vector<int32_t> bid(n, -1);
  std::vector<int32_t> b_out, b_depth, b_cells;
  for (int32_t i = 0; i < n; i++) {
    const int32_t o = outlet[i];
    if (bid[o] < 0) {
      bid[o] = int32_t(b_out.size());
      b_out.push_back(o);
      b_depth.push_back(0);
      b_cells.push_back(0);
    }
    const int32_t b = bid[o];
    b_depth[b] = std::max(b_depth[b], level[i] + 1);
    b_cells[b]++;
  }
  const int32_t nb = int32_t(b_out.size());
  std::vector<int32_t> border(nb);
  for (int32_t b = 0; b < nb; b++) border[b] = b;
  std::sort(border.begin(), border.end(), [&](int32_t a, int32_t b) {
    if (b_depth[a] != b_depth[b]) return b_depth[a] > b_depth[b];
    if (b_cells[a] != b_cells[b]) return b_cells[a] > b_cells[b];
    return b_out[a] < b_out[b];
  });
  // greedy packing of basins into routing groups of about group_cells cells
  std::vector<int32_t> group_of_basin(nb, -1);
  int32_t g = -1, cur_cells = 0;
  std::vector<int32_t> g_depth;
  for (int32_t k = 0; k < nb; k++) {
    const int32_t b = border[k];
    if (g < 0 || cur_cells + b_cells[b] > group_cells) {
      g++;
      cur_cells = 0;
      g_depth.push_back(0);
    }
    group_of_basin[b] = g;
    cur_cells += b_cells[b];
    g_depth[g] = std::max(g_depth[g], b_depth[b]);
  }
  const int32_t n_groups0 = g + 1;
  // XCD-aware placement: bin the groups over the XCDs by cell count (largest
  // first into the lightest bin) and order them bin by bin, so that a bin's cells
  // are contiguous and both kernels can run a group's cells on one XCD.
  std::vector<int32_t> g_cells(n_groups0, 0);
  for (int32_t i = 0; i < n; i++) g_cells[group_of_basin[bid[outlet[i]]]]++;
  std::vector<int32_t> gsort(n_groups0);
  for (int32_t gg = 0; gg < n_groups0; gg++) gsort[gg] = gg;
  std::stable_sort(gsort.begin(), gsort.end(), [&](int32_t a, int32_t b) { return g_cells[a] > g_cells[b]; });
  const int nx = h->n_xcd;
  // Partitions: the groups at least 0.8 times as deep as the deepest form the
  // first when they hold at most 30 % of the cells (and not every group); their
  // days then run beside the others' (WG_PARTITIONS=1 keeps one partition).  Only
  // for up to 8 members: a day of few members is latency-bound, and overlapping
  // the deep groups' walk pays; with many it is bandwidth-bound, and the deep
  // groups' cells, on few XCDs, would load those more than the rest (measured
  // 24 % slower at 256 members).
  constexpr int32_t kPartitionMaxMembers = 8;
  std::vector<int32_t> part_of(n_groups0, 0);
  int n_parts = 1;
  {
    int32_t max_d = 0;
    for (int32_t d : g_depth) max_d = std::max(max_d, d);
    const int32_t deep = (4 * max_d + 4) / 5;
    int64_t deep_cells = 0;
    int32_t n_deep = 0;
    for (int32_t gg = 0; gg < n_groups0; gg++)
      if (g_depth[gg] >= deep) {
        n_deep++;
        deep_cells += g_cells[gg];
      }
    const char *env = getenv("WG_PARTITIONS");
    const bool allow = !(env && atoi(env) == 1);
    if (allow && h->m <= kPartitionMaxMembers && n_deep > 0 && n_deep < n_groups0 &&
        deep_cells * 10 <= int64_t(n) * 3) {
      n_parts = 2;
      for (int32_t gg = 0; gg < n_groups0; gg++) part_of[gg] = g_depth[gg] >= deep ? 0 : 1;
    }
  }
  // XCD bins of each partition, [part][xcd], each partition balanced on its own
  // (balancing the two together measured slower: it crowds the other partition's
  // cells onto six XCDs while the deep groups' cells sit on two)
  std::vector<std::vector<std::vector<int32_t>>> bin_groups(n_parts, std::vector<std::vector<int32_t>>(nx));
  for (int pp = 0; pp < n_parts; pp++) {
    std::vector<int64_t> bin_load(nx, 0);
    for (int32_t gg : gsort) {
      if (part_of[gg] != pp) continue;
      int best = 0;
      for (int x = 1; x < nx; x++)
        if (bin_load[x] < bin_load[best]) best = x;
      bin_groups[pp][best].push_back(gg);
      bin_load[best] += g_cells[gg];
    }
  }
  std::vector<int32_t> new_gid(n_groups0);
  std::vector<int32_t> depth_new;
  for (int pp = 0; pp < n_parts; pp++)
    for (int x = 0; x < nx; x++)
      for (int32_t gg : bin_groups[pp][x]) {
        new_gid[gg] = int32_t(depth_new.size());
        depth_new.push_back(g_depth[gg]);
      }
  h->n_groups = n_groups0;
  h->max_depth = 0;
  for (int32_t d : depth_new) h->max_depth = std::max(h->max_depth, d);
  auto gid = [&](int32_t cell) { return new_gid[group_of_basin[bid[outlet[cell]]]]; };
  // device order: (bin, group, level, cells with a global lake or reservoir last,
  // routing order).  The cells of one level are independent; putting the few with
  // a lake or reservoir together keeps the general routing step, which only they
  // need on every day, to few wavefronts.
  auto lakeres = [&](int32_t c) { return (st->lake_area[c] > 0. || st->reservoir_area[c] > 0.) ? 1 : 0; };
  std::vector<int32_t> cells(n);
  for (int32_t i = 0; i < n; i++) cells[i] = i;
  std::sort(cells.begin(), cells.end(), [&](int32_t a, int32_t b) {
    const int32_t ga = gid(a), gb = gid(b);
    if (ga != gb) return ga < gb;
    if (level[a] != level[b]) return level[a] < level[b];
    if (lakeres(a) != lakeres(b)) return lakeres(a) < lakeres(b);
    return st->rout_order[a] < st->rout_order[b];
  });
  h->perm = cells;
  h->inv.assign(n, 0);
  for (int32_t p = 0; p < n; p++) h->inv[cells[p]] = p;
  // level slots
  h->group_first_level.assign(h->n_groups + 1, 0);
  for (int32_t gg = 0; gg < h->n_groups; gg++) h->group_first_level[gg + 1] = h->group_first_level[gg] + depth_new[gg];
  const int32_t n_slots = h->group_first_level[h->n_groups];
  h->level_off.assign(n_slots + 1, n);
  std::vector<int32_t> slot_count(n_slots, 0);
  for (int32_t i = 0; i < n; i++) slot_count[h->group_first_level[gid(i)] + level[i]]++;
  int32_t acc = 0;
  for (int32_t s = 0; s < n_slots; s++) {
    h->level_off[s] = acc;
    acc += slot_count[s];
  }
  h->level_off[n_slots] = acc;
  // block tables: within each partition, block 8j + x runs on XCD x
  h->parts.assign(n_parts, WgPart{});
  h->block_group.clear();
  {
    int32_t gg = 0;
    for (int pp = 0; pp < n_parts; pp++) {
      WgPart &P = h->parts[pp];
      P.bin_start.assign(nx, 0);
      P.bin_end.assign(nx, 0);
      const int32_t g0 = gg;
      int32_t max_groups = 1;
      for (int x = 0; x < nx; x++) {
        P.bin_start[x] = h->level_off[h->group_first_level[gg]];
        gg += int32_t(bin_groups[pp][x].size());
        P.bin_end[x] = h->level_off[h->group_first_level[gg]];
        max_groups = std::max(max_groups, int32_t(bin_groups[pp][x].size()));
      }
      P.cell_begin = h->level_off[h->group_first_level[g0]];
      P.cell_end = h->level_off[h->group_first_level[gg]];
      P.group_block_off = int32_t(h->block_group.size());
      P.n_group_blocks = nx * max_groups;
      h->block_group.resize(size_t(P.group_block_off) + P.n_group_blocks, -1);
      int32_t g1 = g0;
      for (int x = 0; x < nx; x++)
        for (size_t j = 0; j < bin_groups[pp][x].size(); j++)
          h->block_group[size_t(P.group_block_off) + j * nx + x] = g1++;
    }
  }
  build_tiles(h);
  return WG_OK;
}

// The wave walk's batch table for W wavefronts: the batches of the wide levels of
// every group dealt round-robin, each flagged when one of its cells has a global
// lake or reservoir (from the chain links' flags).
void wave_tables(const wg_handle *H, int nw, std::vector<int32_t> &item_off, std::vector<WgWaveItem> &items) {
  item_off.assign(size_t(H->n_groups) * nw + 1, 0);
  items.clear();
  for (int32_t gg = 0; gg < H->n_groups; gg++) {
    const int32_t l0 = H->group_first_level[gg];
    for (int w = 0; w < nw; w++) {
      item_off[size_t(gg) * nw + w] = int32_t(items.size());
      for (int32_t L = 1; L <= H->wave_wide[gg]; L++) {
        const int32_t a = H->level_off[l0 + L], e = H->level_off[l0 + L + 1];
        int32_t j = 0;
        for (int32_t st0 = a; st0 < e; st0 += 64, j++) {
          if (j % nw != w) continue;
          const int32_t end = std::min(st0 + 64, e);
          int32_t lr = 0;
          for (int32_t q = st0; q < end; q++) lr |= (H->clink[q].flags & (CFL_LAKE | CFL_RES)) ? 1 : 0;
          items.push_back(WgWaveItem{st0, end, L, lr});
        }
      }
    }
  }
  item_off[size_t(H->n_groups) * nw] = int32_t(items.size());
  if (items.empty()) items.push_back(WgWaveItem{0, 0, 0, 0});
}

int upload_tiles(wg_handle *h) {
  if (int32_t(h->tile_start.size()) > h->tile_capacity) return fail(h, WG_E_ARG, "tile table too large");
  HIP_CHECK(h, hipMemcpy(const_cast<int32_t *>(h->S.tile_start), h->tile_start.data(),
                         h->tile_start.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(h, hipMemcpy(const_cast<int32_t *>(h->S.tile_end), h->tile_end.data(),
                         h->tile_end.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  return WG_OK;
}

}  // namespace

extern "C" int wg_create(int32_t n, const wg_static_fields *st, const wg_options_profile *opt, int32_t n_members,
                         wg_handle **out) {
  if (!out) return WG_E_ARG;
  *out = nullptr;
  // rows of fewer than 2^27 cells: the chain kernels address a member's row of
  // doubles or 16-byte pairs with 32-bit byte offsets below 2^31
  if (n <= 0 || n >= (1 << 27) || !st || !opt || n_members <= 0 || n_members > 65535) return WG_E_ARG;
  const void *req[] = {st->row, st->cell_area, st->contfreq, st->aindex, st->ldd, st->land_cover, st->built_up,
                       st->smax, st->gw_factor, st->rgmax, st->texture, st->elevation, st->params, st->glo_lake,
                       st->loc_lake, st->glo_wetland, st->loc_wetland, st->lake_area, st->reservoir_area,
                       st->glo_res, st->stor_cap, st->mean_outflow, st->res_type, st->start_month,
                       st->river_length, st->river_slope, st->roughness, st->bankfull_flow, st->downstream,
                       st->rout_order, st->lct};
  for (const void *p : req)
    if (!p) return WG_E_ARG;
  std::unique_ptr<wg_handle> h(new wg_handle());
  h->n = n;
  h->m = n_members;
  if (!options_supported(opt)) return WG_E_ARG;
  h->subtract_use = opt->subtract_use;
  for (int32_t i = 0; i < n; i++) {
    if (st->aindex[i] != 0 && st->aindex[i] != 1) return WG_E_INPUT;  // daily.cpp:345-347
    if (st->land_cover[i] < 1 || st->land_cover[i] > WG_NLCT) return WG_E_INPUT;
    if (st->reservoir_area[i] > 0. && !(st->res_type[i] == 1 || st->res_type[i] == 2)) return WG_E_INPUT;
    if (st->reservoir_area[i] > 0. && (st->start_month[i] < 1 || st->start_month[i] > 12)) return WG_E_INPUT;
  }
  hipError_t e = hipGetDevice(&h->device);
  if (e != hipSuccess) return WG_E_HIP;
  e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return WG_E_HIP;
  // the wave walk on eight wavefronts for few members (its wide levels finish
  // sooner), on four for many (more workgroups fit per CU)
  h->block_threads_chain = n_members <= 4 ? 512 : 256;
  if (const char *v = getenv("WG_CHAIN_THREADS")) h->block_threads_chain = std::max(64, std::min(512, atoi(v)));
  if (const char *v = getenv("WG_XCDS")) h->n_xcd = std::max(1, std::min(64, atoi(v)));
  if (const char *v = getenv("WG_CHAIN_MODE")) h->chain_mode = std::max(0, std::min(3, atoi(v)));
  if (h->chain_mode != 3) h->block_threads_chain = std::min(h->block_threads_chain, 256);
  int group_cells = 1024;
  if (const char *v = getenv("WG_GROUP_CELLS")) group_cells = std::max(1, atoi(v));
  int rc = build_schedule(h.get(), st, group_cells);
  if (rc) {
    *out = h.release();
    return rc;
  }
  if (h->max_depth > WG_MAX_GROUP_DEPTH && h->chain_mode == 1) h->chain_mode = 0;  // dataflow keeps levels in LDS
  wg_handle *H = h.get();
  const int32_t M = n_members;
  {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, H->device) != hipSuccess)
      lds = 65536;
    int32_t biggest = 0;
    for (int32_t gg = 0; gg < H->n_groups; gg++)
      biggest = std::max(biggest, H->level_off[H->group_first_level[gg + 1]] - H->level_off[H->group_first_level[gg]]);
    // a routing group's transported volumes (8 B per cell) and chain flags (1 B)
    // live in LDS while the chain kernels walk it
    if ((size_t(biggest) * 9 + 15) / 16 * 16 > size_t(lds)) {
      *out = h.release();
      return fail(*out, WG_E_INPUT, "a routing group of %d cells exceeds the %d bytes of LDS per workgroup "
                  "(largest basin too large)", biggest, lds);
    }
  }
  const double *par = st->params;
  auto P = [&](int id, int32_t i) { return par[size_t(id) * n + i]; };

  // tables (lai.cpp:86-89; daily.cpp:1925-1959)
  for (int c = 0; c < WG_NLCT; c++) {
    H->T.albedo[c] = st->lct->albedo[c];
    H->T.albedo_snow[c] = st->lct->albedo_snow[c];
    H->T.ddf[c] = st->lct->ddf[c];
    H->T.emissivity[c] = st->lct->emissivity[c];
    H->T.lai_a[c] = float(0.1 * st->lct->decid_frac[c]);
    H->T.lai_b[c] = (1 - st->lct->decid_frac[c]) * st->lct->evergreen_red[c];
    H->T.initial_days[c] = st->lct->initial_days[c];
  }

  // derived statics in reference order
  std::vector<double> alpha(n), river_length(n), river_slope(n), roughness(n), width_bf(n), bottom(n),
      stor_max(n), lake_depth(n), wetl_depth(n), f_glo_lake(n), fswb_init(n), k_sw(n), e_sw(n), inv_ksw(n),
      ome_sw(n), k_gw(n), e_gw(n), inv_kgw(n), ome_gw(n), vel_c1(n), vel_c2(n), rw_c0(n), rw_c1(n), rw_c2(n),
      rw_c3(n);
  std::vector<float> lai_max(n);
  std::vector<int8_t> contcell(n);
  for (int32_t i = 0; i < n; i++) {
    alpha[i] = st->aindex[i] == 1 ? P(WG_P_PTC_ARI, i) : P(WG_P_PTC_HUM, i);
    lai_max[i] = float(P(WG_M_LAI, i) * st->lct->lai[st->land_cover[i] - 1]);
    contcell[i] = (st->contfreq[i] < 0.000001) ? 0 : 1;
    river_length[i] = double(st->river_length[i]);
    river_length[i] *= st->contfreq[i] / 100.;
    river_slope[i] = double(st->river_slope[i]);
    roughness[i] = double(st->roughness[i]);
    double bf = double(st->bankfull_flow[i]);
    if (bf < 0.05) bf = 0.05;
    width_bf[i] = 2.71 * wg_pow(bf, 0.557);
    const double depth_bf = 0.349 * wg_pow(bf, 0.341);
    bottom[i] = width_bf[i] - 2.0 * 2.0 * depth_bf;
    stor_max[i] = river_length[i] * 0.5 * depth_bf / 1000. * (bottom[i] / 1000. + width_bf[i] / 1000.);
    lake_depth[i] = (P(WG_P_LAK_D, i) * 0.001);
    wetl_depth[i] = (P(WG_P_WET_D, i) * 0.001);
    const double fll = st->loc_lake[i] / 100., flw = st->loc_wetland[i] / 100., fgw = st->glo_wetland[i] / 100.;
    f_glo_lake[i] = st->glo_lake[i] / 100.;
    fswb_init[i] = fll + flw + fgw;
    k_sw[i] = P(WG_P_SWOUTF_C, i);
    e_sw[i] = wg_exp(-1. * k_sw[i]);
    inv_ksw[i] = 1. / k_sw[i];
    ome_sw[i] = 1. - wg_exp(-1. * k_sw[i]);
    k_gw[i] = P(WG_P_GWOUTF_C, i);
    e_gw[i] = wg_exp(-1. * k_gw[i]);
    inv_kgw[i] = 1. / k_gw[i];
    ome_gw[i] = 1. - wg_exp(-1. * k_gw[i]);
    // constants of getNewRiverVelocity (routing.cpp:7309-7341) and of the river width
    // (routing.cpp:3546-3560), in the reference's operation order
    vel_c1[i] = 1. / (P(WG_M_RIVRGH_C, i) * roughness[i]);
    vel_c2[i] = wg_pow(river_slope[i], 0.5);
    rw_c0[i] = -bottom[i] / (4. * 1000.);
    rw_c1[i] = bottom[i] / 1000. * bottom[i] / (16. * 1000.);
    rw_c2[i] = bottom[i] / 1000.;
    rw_c3[i] = width_bf[i] / 1000.;
  }

  const std::vector<double> prov_rel = reservoir_prov_rel(st, n);
  H->glo_res_prev.assign(st->glo_res, st->glo_res + n);  // routing.cpp:1085-1088

  // initial state (one member, reference order)
  std::vector<double> s_river(n), s_loc_lake(n), s_loc_wetl(n), s_glo_lake(n), s_glo_wetl(n), s_res(n),
      s_red_river(n), s_rfn(n), s_laf(n), s_one(n, 1.0), s_krel(n, 0.1);
  for (int32_t i = 0; i < n; i++) {
    s_river[i] = stor_max[i];
    s_loc_lake[i] = (st->loc_lake[i] / 100.) * st->cell_area[i] * lake_depth[i];
    s_loc_wetl[i] = (st->loc_wetland[i] / 100.) * st->cell_area[i] * wetl_depth[i];
    s_glo_lake[i] = st->lake_area[i] * lake_depth[i];
    s_glo_wetl[i] = (st->glo_wetland[i] / 100.) * st->cell_area[i] * wetl_depth[i];
    s_res[i] = st->stor_cap[i];
    s_red_river[i] = 0.5;
    s_rfn[i] = s_red_river[i] * river_length[i] * width_bf[i] / 1000. / st->cell_area[i];
    s_laf[i] = (st->contfreq[i] -
                (st->glo_lake[i] + st->glo_wetland[i] + st->loc_lake[i] + st->loc_wetland[i] + st->glo_res[i]));
    if (s_laf[i] < 0.) s_laf[i] = 0.;
  }

  // ---- upload statics ----
  WgStatic &S = H->S;
  S.n = n;
  S.n_members = M;
#define UP(field, vec)                          \
  do {                                          \
    auto v_ = permute(H, (vec).data());         \
    int r_ = upload(H, &field, v_);             \
    if (r_) { *out = h.release(); return r_; }  \
  } while (0)
#define UPR(field, ptr_)                        \
  do {                                          \
    auto v_ = permute(H, (ptr_));               \
    int r_ = upload(H, &field, v_);             \
    if (r_) { *out = h.release(); return r_; }  \
  } while (0)
  {
    std::vector<int32_t> ref(n);
    for (int32_t i = 0; i < n; i++) ref[i] = i;
    int32_t *d = nullptr;
    UP(d, ref);
    S.ref_index = d;
  }
  {
    double *d;
    UPR(d, st->cell_area); S.cell_area = d;
    UPR(d, st->contfreq); S.contfreq = d;
    UP(d, alpha); S.alpha = d;
    std::vector<double> tmp(n);
    const int pids[] = {WG_P_PET_MXDY, WG_M_PREC, WG_M_NETRAD, WG_P_MCWH, WG_P_T_SNOWFZ, WG_P_T_SNOWMT,
                        WG_P_T_GRADNT, WG_M_DEGDAY_F, WG_P_PCRITGWA, WG_M_EVAREDEX, WG_M_RIVRGH_C};
    const double **dst[] = {&S.pet_mxdy, &S.m_prec, &S.m_netrad, &S.mcwh, &S.t_snowfz, &S.t_snowmt,
                            &S.t_gradnt, &S.degday, &S.pcritgwa, &S.evarexp, &S.m_rivrgh};
    for (size_t q = 0; q < sizeof(pids) / sizeof(pids[0]); q++) {
      for (int32_t i = 0; i < n; i++) tmp[i] = P(pids[q], i);
      UP(d, tmp);
      *dst[q] = d;
    }
    UPR(d, st->loc_lake); S.loc_lake = d;
    UPR(d, st->loc_wetland); S.loc_wetland = d;
    UPR(d, st->glo_wetland); S.glo_wetland = d;
    UPR(d, st->lake_area); S.lake_area = d;
    UPR(d, st->reservoir_area); S.reservoir_area = d;
    UPR(d, st->stor_cap); S.stor_cap = d;
    UPR(d, st->mean_outflow); S.mean_outflow = d;
    UP(d, prov_rel); S.res_prov_rel = d;
    UP(d, river_length); S.river_length = d;
    UP(d, river_slope); S.river_slope = d;
    UP(d, vel_c1); S.vel_c1 = d;
    UP(d, vel_c2); S.vel_c2 = d;
    UP(d, rw_c0); S.rw_c0 = d;
    UP(d, rw_c1); S.rw_c1 = d;
    UP(d, rw_c2); S.rw_c2 = d;
    UP(d, rw_c3); S.rw_c3 = d;
    UP(d, roughness); S.roughness = d;
    UP(d, width_bf); S.width_bf = d;
    UP(d, bottom); S.bottom_width = d;
    UP(d, stor_max); S.river_stor_max = d;
    UP(d, lake_depth); S.lake_depth = d;
    UP(d, wetl_depth); S.wetl_depth = d;
    UP(d, f_glo_lake); S.f_glo_lake = d;
    UP(d, fswb_init); S.fswb_init = d;
    UP(d, k_sw); S.k_sw = d;
    UP(d, e_sw); S.e_sw = d;
    UP(d, inv_ksw); S.inv_ksw = d;
    UP(d, ome_sw); S.ome_sw = d;
    UP(d, k_gw); S.k_gw = d;
    UP(d, e_gw); S.e_gw = d;
    UP(d, inv_kgw); S.inv_kgw = d;
    UP(d, ome_gw); S.ome_gw = d;
    std::vector<double> cf100(n);
    for (int32_t i = 0; i < n; i++) cf100[i] = st->contfreq[i] / 100.;
    UP(d, cf100); S.cf100 = d;
  }
  {
    int8_t *d8;
    UPR(d8, st->ldd); S.ldd = d8;
    UPR(d8, st->land_cover); S.land_cover = d8;
    UPR(d8, st->texture); S.texture = d8;
    UP(d8, contcell); S.contcell = d8;
    std::vector<int8_t> cflags(n);
    for (int32_t i = 0; i < n; i++) {
      int f = 0;
      if (st->lake_area[i] > 0.) f |= CFL_LAKE;
      if (st->reservoir_area[i] > 0.) f |= CFL_RES;
      if (st->glo_wetland[i] > 0) f |= CFL_GWET;
      if (st->aindex[i] == 1 && st->ldd[i] >= 0) f |= CFL_ARID;
      if (st->ldd[i] >= 0) f |= CFL_LDD;
      if (contcell[i]) f |= CFL_CONT;
      cflags[i] = int8_t(f);
    }
    UP(d8, cflags); S.cflags = d8;
    UPR(d8, st->res_type); S.res_type = d8;
    UPR(d8, st->start_month); S.start_month = d8;
    int16_t *d16;
    UPR(d16, st->aindex); S.aindex = d16;
    UPR(d16, st->rgmax); S.rgmax = d16;
    float *df;
    UPR(df, st->built_up); S.built_up = df;
    UPR(df, st->gw_factor); S.gw_factor = df;
    UP(df, lai_max); S.lai_max = df;
    // elevation: channel 0 [N], sub-cells as packed int16 pairs [50][N]
    std::vector<int16_t> e0(n);
    std::vector<int32_t> pairs(size_t(WG_DEV_NPAIR) * n);
    for (int32_t p = 0; p < n; p++) {
      const int16_t *ev = st->elevation + size_t(H->perm[p]) * WG_NELEV;
      e0[p] = ev[0];
      for (int k = 0; k < WG_DEV_NPAIR; k++)
        pairs[size_t(k) * n + p] = int32_t(uint32_t(uint16_t(ev[1 + 2 * k])) | (uint32_t(uint16_t(ev[2 + 2 * k])) << 16));
    }
    int16_t *de0;
    int r_ = upload(H, &de0, e0);
    if (r_) { *out = h.release(); return r_; }
    S.elev0 = de0;
    int32_t *dpairs;
    r_ = upload(H, &dpairs, pairs);
    if (r_) { *out = h.release(); return r_; }
    S.elev_pairs = dpairs;
  }
  // upstream CSR in device indices, each list in routing order (the order of the
  // reference's serial += into the downstream cell's inflow)
  {
    std::vector<std::vector<int32_t>> ups(n);
    for (int32_t i = 0; i < n; i++) {
      const int32_t d = st->downstream[i];
      if (d > 0) ups[H->inv[d - 1]].push_back(H->inv[i]);
    }
    // group of every device cell, for group-local upstream indices (the chain
    // kernel keeps a group's transported volumes in LDS)
    std::vector<int32_t> gstart(n, 0);
    H->max_group_cells = 0;
    for (int32_t gg = 0; gg < H->n_groups; gg++) {
      const int32_t a = H->level_off[H->group_first_level[gg]], b = H->level_off[H->group_first_level[gg + 1]];
      for (int32_t p = a; p < b; p++) gstart[p] = a;
      H->max_group_cells = std::max(H->max_group_cells, b - a);
    }

    std::vector<int32_t> off(n + 1, 0), idx;
    for (int32_t p = 0; p < n; p++)
      if (ups[p].size() > 8) {
        *out = h.release();
        return fail(*out, WG_E_INPUT, "cell %d has %d upstream cells (D8 allows 8)", H->perm[p],
                    int(ups[p].size()));
      }
    for (int32_t p = 0; p < n; p++) {
      std::sort(ups[p].begin(), ups[p].end(),
                [&](int32_t a, int32_t b) { return st->rout_order[H->perm[a]] < st->rout_order[H->perm[b]]; });
      off[p + 1] = off[p] + int32_t(ups[p].size());
      for (int32_t u : ups[p]) idx.push_back(u - gstart[p]);
    }
    int32_t *d;
    int r_ = upload(H, &d, off);
    if (r_) { *out = h.release(); return r_; }
    S.up_off = d;
    r_ = upload(H, &d, idx);
    if (r_) { *out = h.release(); return r_; }
    S.up_idx = d;
    std::vector<int32_t> none(n, -1);
    r_ = upload(H, &d, none);
    if (r_) { *out = h.release(); return r_; }
    S.station = d;
    // routing links and outflow constants of the chain storages
    H->clink.assign(n, WgChainLink{});
    std::vector<WgChainStatic> cst(n);
    // Kahn level of every device cell counted from its group's first level
    std::vector<int32_t> rel_level(n, 0);
    for (int32_t gg = 0; gg < H->n_groups; gg++)
      for (int32_t l = H->group_first_level[gg]; l < H->group_first_level[gg + 1]; l++)
        for (int32_t q = H->level_off[l]; q < H->level_off[l + 1]; q++) rel_level[q] = l - H->group_first_level[gg];
    for (int32_t p = 0; p < n; p++) {
      const int32_t i = H->perm[p];
      WgChainStatic &c = cst[p];
      c.inv_ksw = inv_ksw[i];
      c.ome_sw = ome_sw[i];
      c.inv_kgw = inv_kgw[i];
      c.ome_gw = ome_gw[i];
      c.area = st->cell_area[i];
      c.cf100 = st->contfreq[i] / 100.;
      int f = 0;
      if (st->lake_area[i] > 0.) f |= CFL_LAKE;
      if (st->reservoir_area[i] > 0.) f |= CFL_RES;
      if (st->glo_wetland[i] > 0) f |= CFL_GWET;
      if (st->aindex[i] == 1 && st->ldd[i] >= 0) f |= CFL_ARID;
      if (st->ldd[i] >= 0) f |= CFL_LDD;
      if (contcell[i]) f |= CFL_CONT;
      if (st->reservoir_area[i] > 0.) f |= (int(st->res_type[i]) & 0xff) << 8 | (int(st->start_month[i]) & 0xff) << 16;
      WgChainLink &lk = H->clink[p];
      lk.flags = f;
      lk.station = -1;
      lk.nu = int16_t(off[p + 1] - off[p]);
      lk.level = rel_level[p];
      for (int u = 0; u < 8; u++) lk.up[u] = int16_t(u < lk.nu ? idx[off[p] + u] : 0);
    }
    // wave walk: the leading wide levels (more than 64 cells) of each group above
    // level 0, in level-aligned batches of 64 cells dealt round-robin to the
    // wavefronts of the group's workgroup
    {
      std::vector<int32_t> wide(size_t(H->n_groups), 0);
      for (int32_t gg = 0; gg < H->n_groups; gg++) {
        const int32_t l0 = H->group_first_level[gg], l1 = H->group_first_level[gg + 1];
        int32_t la = 0;
        while (l0 + 1 + la < l1 && H->level_off[l0 + 2 + la] - H->level_off[l0 + 1 + la] > 64) la++;
        wide[gg] = la;
      }
      r_ = upload(H, &H->d_wave_wide, wide);
      if (r_) { *out = h.release(); return r_; }
      H->wave_wide = wide;
      for (int t = 0; t < 2; t++) {
        std::vector<int32_t> item_off;
        std::vector<WgWaveItem> items;
        wave_tables(H, t == 0 ? 4 : 8, item_off, items);
        r_ = upload(H, &H->d_wave_item_off[t], item_off);
        if (r_) { *out = h.release(); return r_; }
        r_ = upload(H, &H->d_wave_items[t], items);
        if (r_) { *out = h.release(); return r_; }
      }
    }
    // the largest upstream count of each level
    for (int32_t l = 0; l + 1 < int32_t(H->level_off.size()); l++) {
      int16_t mx = 0;
      for (int32_t q = H->level_off[l]; q < H->level_off[l + 1]; q++) mx = std::max(mx, H->clink[q].nu);
      for (int32_t q = H->level_off[l]; q < H->level_off[l + 1]; q++) H->clink[q].nu_level = mx;
    }
    WgI4 *dlk;
    r_ = upload(H, &dlk, pack_links(H->clink));
    if (r_) { *out = h.release(); return r_; }
    S.clink = dlk;
    std::vector<WgD2> cst_p(size_t(3) * n);
    for (int32_t p = 0; p < n; p++) {
      const WgChainStatic &c = cst[p];
      cst_p[p] = WgD2{c.inv_ksw, c.ome_sw};
      cst_p[size_t(n) + p] = WgD2{c.inv_kgw, c.ome_gw};
      cst_p[size_t(2) * n + p] = WgD2{c.area, c.cf100};
    }
    WgD2 *dcs;
    r_ = upload(H, &dcs, cst_p);
    if (r_) { *out = h.release(); return r_; }
    S.cstat = dcs;
    // tile tables sized for the smallest cell block the tuner tries (64 threads)
    int32_t *dt;
    {
      H->tile_capacity = 0;
      for (const WgPart &P : H->parts) {
        int32_t cap = 1;
        for (int x = 0; x < H->n_xcd; x++) cap = std::max(cap, (P.bin_end[x] - P.bin_start[x] + 63) / 64);
        H->tile_capacity += cap * H->n_xcd;
      }
    }
    r_ = dev_alloc(H, &dt, size_t(H->tile_capacity));
    if (r_) { *out = h.release(); return r_; }
    S.tile_start = dt;
    r_ = dev_alloc(H, &dt, size_t(H->tile_capacity));
    if (r_) { *out = h.release(); return r_; }
    S.tile_end = dt;
    r_ = upload_tiles(H);
    if (r_) { *out = h.release(); return r_; }
    r_ = upload(H, &dt, H->block_group);
    if (r_) { *out = h.release(); return r_; }
    S.block_group = dt;
    r_ = upload(H, &H->d_level_off, H->level_off);
    if (r_) { *out = h.release(); return r_; }
    r_ = upload(H, &H->d_group_first, H->group_first_level);
    if (r_) { *out = h.release(); return r_; }
  }
#undef UP
#undef UPR

  // ---- per-member state ----
  WgState &X = H->X;
  const size_t NM = size_t(n) * M;
  double **dstate[] = {&X.canopy, &X.soil, &X.snow, &X.transfer, &X.lai_precsum, &X.gw, &X.loc_lake, &X.loc_wetl,
                       &X.glo_lake, &X.glo_wetl, &X.res, &X.river, &X.red_ll, &X.red_lw, &X.red_gl, &X.red_gw,
                       &X.red_res, &X.red_river, &X.river_frac_next, &X.river_frac_change, &X.fswb_next, &X.laf,
                       &X.laf_prev, &X.k_release, &X.gamma, &X.cfa, &X.tv, &X.d_surf,
                       &X.d_recharge, &X.d_ow_pet, &X.d_aet, &X.d_prec};
  for (double **p : dstate) {
    rc = dev_alloc(H, p, NM);
    if (rc) { *out = h.release(); return rc; }
  }
  rc = dev_alloc(H, &X.crec, NM * (WG_DEV_NREC / 2));
  if (!rc) rc = dev_alloc(H, &X.lai_days, NM);
  if (!rc) rc = dev_alloc(H, &X.lai_status, NM);
  if (!rc) rc = dev_alloc(H, &X.smax, NM);
  if (!rc) rc = dev_alloc(H, &X.snow2, NM * WG_DEV_NPAIR);
  if (!rc) rc = dev_alloc(H, &X.snow_any, NM);
  if (!rc) rc = dev_alloc(H, &X.flow_error, 1);
  if (!rc) rc = dev_alloc(H, &X.st_sum, 1);
  if (!rc) rc = dev_alloc(H, &X.st_daily, 1);
  if (!rc) rc = dev_alloc(H, &H->d_clock, 1);
  if (!rc) rc = dev_alloc(H, &H->d_seeds, size_t(M));
  if (rc) { *out = h.release(); return rc; }
  X.member_seed = H->d_seeds;
  X.perturb = 0;
  H->snow_ch0.assign(NM, 0.0);
  auto put = [&](double *dst, const std::vector<double> &ref_vec) -> int {
    auto v = replicate(permute(H, ref_vec.data()), M);
    hipError_t e2 = hipMemcpy(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice);
    return e2 == hipSuccess ? WG_OK : fail(H, WG_E_HIP, "hipMemcpy: %s", hipGetErrorString(e2));
  };
  std::vector<double> gam(n), cfa(n);
  for (int32_t i = 0; i < n; i++) {
    gam[i] = P(WG_P_GAMRUN_C, i);
    cfa[i] = P(WG_P_CFA, i);
  }
  rc = put(X.river, s_river);
  if (!rc) rc = put(X.loc_lake, s_loc_lake);
  if (!rc) rc = put(X.loc_wetl, s_loc_wetl);
  if (!rc) rc = put(X.glo_lake, s_glo_lake);
  if (!rc) rc = put(X.glo_wetl, s_glo_wetl);
  if (!rc) rc = put(X.res, s_res);
  if (!rc) rc = put(X.red_ll, s_one);
  if (!rc) rc = put(X.red_lw, s_one);
  if (!rc) rc = put(X.red_gl, s_one);
  if (!rc) rc = put(X.red_gw, s_one);
  if (!rc) rc = put(X.red_res, s_one);
  if (!rc) rc = put(X.red_river, s_red_river);
  if (!rc) rc = put(X.river_frac_next, s_rfn);
  if (!rc) rc = put(X.fswb_next, fswb_init);
  if (!rc) rc = put(X.laf, s_laf);
  if (!rc) rc = put(X.k_release, s_krel);
  if (!rc) rc = put(X.gamma, gam);
  if (!rc) rc = put(X.cfa, cfa);
  if (!rc) {
    auto v = replicate(permute(H, st->smax), M);
    hipError_t e2 = hipMemcpy(X.smax, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e2 != hipSuccess) rc = fail(H, WG_E_HIP, "hipMemcpy: %s", hipGetErrorString(e2));
  }
  if (rc) { *out = h.release(); return rc; }
  e = hipEventCreate(&H->ev0);
  if (e == hipSuccess) e = hipEventCreate(&H->ev1);
  if (e == hipSuccess && H->parts.size() > 1) {
    // the partitions' streams (default priority) and their fork/join events
    e = hipEventCreateWithFlags(&H->ev_fork, hipEventDisableTiming);
    for (WgPart &P : H->parts) {
      if (e == hipSuccess) e = hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&P.done, hipEventDisableTiming);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(H->stream);
  if (e != hipSuccess) {
    *out = h.release();
    return WG_E_HIP;
  }
  *out = h.release();
  return WG_OK;
}

extern "C" int wg_destroy(wg_handle *h) {
  if (!h) return WG_E_ARG;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (WgPart &P : h->parts) {
    if (P.stream) (void)hipStreamSynchronize(P.stream);
    for (auto &kv : P.graphs) (void)hipGraphExecDestroy(kv.second);
    if (P.done) (void)hipEventDestroy(P.done);
    if (P.stream) (void)hipStreamDestroy(P.stream);
  }
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto &b : h->bufs) (void)hipFree(b.p);
  if (h->d_forcing) (void)hipFree(h->d_forcing);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return WG_OK;
}

extern "C" int wg_get_state(wg_handle *h, int32_t member, wg_state *o) {
  if (!h || !o || member < 0 || member >= h->m) return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  const int32_t n = h->n;
  const size_t off = size_t(member) * n;
  std::vector<double> bd(n);
  std::vector<int32_t> bi(n);
  for (const StateMap &f : kStateMap) {
    if (f.ref_d) {
      double *dst = o->*(f.ref_d);
      if (!dst) continue;
      HIP_CHECK(h, hipMemcpy(bd.data(), (h->X.*(f.dev_d)) + off, n * sizeof(double), hipMemcpyDeviceToHost));
      for (int32_t p = 0; p < n; p++) dst[h->perm[p]] = bd[p];
    } else {
      int32_t *dst = o->*(f.ref_i);
      if (!dst) continue;
      HIP_CHECK(h, hipMemcpy(bi.data(), (h->X.*(f.dev_i)) + off, n * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (int32_t p = 0; p < n; p++) dst[h->perm[p]] = bi[p];
    }
  }
  if (o->snow_elev) {
    std::vector<WgD2> se(size_t(n) * WG_DEV_NPAIR);
    HIP_CHECK(h, hipMemcpy(se.data(), h->X.snow2 + off * WG_DEV_NPAIR, se.size() * sizeof(WgD2),
                           hipMemcpyDeviceToHost));
    for (int32_t p = 0; p < n; p++) {
      const int32_t i = h->perm[p];
      double *dst = o->snow_elev + size_t(i) * WG_NELEV;
      dst[0] = h->snow_ch0[off + i];
      for (int k = 0; k < WG_DEV_NPAIR; k++) {
        dst[1 + 2 * k] = se[size_t(k) * n + p].x;
        dst[2 + 2 * k] = se[size_t(k) * n + p].y;
      }
    }
  }
  o->started = h->started ? 1 : 0;
  return WG_OK;
}

extern "C" int wg_set_state(wg_handle *h, int32_t member, const wg_state *in) {
  if (!h || !in || member < -1 || member >= h->m) return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  const int32_t n = h->n;
  const int32_t m0 = member < 0 ? 0 : member, m1 = member < 0 ? h->m : member + 1;
  std::vector<double> bd(n);
  std::vector<int32_t> bi(n);
  for (const StateMap &f : kStateMap) {
    if (f.ref_d) {
      const double *src = in->*(f.ref_d);
      if (!src) continue;
      for (int32_t p = 0; p < n; p++) bd[p] = src[h->perm[p]];
      for (int32_t mm = m0; mm < m1; mm++)
        HIP_CHECK(h, hipMemcpy((h->X.*(f.dev_d)) + size_t(mm) * n, bd.data(), n * sizeof(double),
                               hipMemcpyHostToDevice));
    } else {
      const int32_t *src = in->*(f.ref_i);
      if (!src) continue;
      for (int32_t p = 0; p < n; p++) bi[p] = src[h->perm[p]];
      for (int32_t mm = m0; mm < m1; mm++)
        HIP_CHECK(h, hipMemcpy((h->X.*(f.dev_i)) + size_t(mm) * n, bi.data(), n * sizeof(int32_t),
                               hipMemcpyHostToDevice));
    }
  }
  if (in->snow_elev) {
    std::vector<WgD2> se(size_t(n) * WG_DEV_NPAIR);
    for (int32_t p = 0; p < n; p++) {
      const double *src = in->snow_elev + size_t(h->perm[p]) * WG_NELEV;
      for (int k = 0; k < WG_DEV_NPAIR; k++) se[size_t(k) * n + p] = WgD2{src[1 + 2 * k], src[2 + 2 * k]};
    }
    // sub-cell storages that are all exactly zero need not be read by the kernel
    std::vector<uint8_t> any(n, 0);
    for (int32_t p = 0; p < n; p++)
      for (int k = 0; k < WG_DEV_NPAIR; k++)
        if (se[size_t(k) * n + p].x != 0. || se[size_t(k) * n + p].y != 0.) any[p] = 1;
    for (int32_t mm = m0; mm < m1; mm++) {
      for (int32_t i = 0; i < n; i++) h->snow_ch0[size_t(mm) * n + i] = in->snow_elev[size_t(i) * WG_NELEV];
      HIP_CHECK(h, hipMemcpy(h->X.snow2 + size_t(mm) * n * WG_DEV_NPAIR, se.data(), se.size() * sizeof(WgD2),
                             hipMemcpyHostToDevice));
      HIP_CHECK(h, hipMemcpy(h->X.snow_any + size_t(mm) * n, any.data(), n, hipMemcpyHostToDevice));
    }
  }
  h->started = in->started != 0;
  return WG_OK;
}

extern "C" int wg_set_member_params(wg_handle *h, const double *gamma, const double *cfa, const float *smax) {
  if (!h) return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  const int32_t n = h->n;
  for (int32_t mm = 0; mm < h->m; mm++) {
    const size_t off = size_t(mm) * n;
    if (gamma) {
      auto v = permute(h, gamma + off);
      HIP_CHECK(h, hipMemcpy(h->X.gamma + off, v.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    if (cfa) {
      auto v = permute(h, cfa + off);
      HIP_CHECK(h, hipMemcpy(h->X.cfa + off, v.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    if (smax) {
      auto v = permute(h, smax + off);
      HIP_CHECK(h, hipMemcpy(h->X.smax + off, v.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  return WG_OK;
}

extern "C" int wg_set_forcing_perturbation(wg_handle *h, const uint64_t *seeds, double prec_sigma,
                                           double temp_sigma) {
  if (!h) return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  if (!seeds) {
    h->X.perturb = 0;
  } else {
    HIP_CHECK(h, hipMemcpy(h->d_seeds, seeds, size_t(h->m) * sizeof(uint64_t), hipMemcpyHostToDevice));
    h->X.perturb = 1;
    h->X.prec_sigma = prec_sigma;
    h->X.temp_sigma = temp_sigma;
  }
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  return WG_OK;
}

namespace {
void clear_graphs(wg_handle *h) {
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  for (WgPart &P : h->parts) {
    for (auto &kv : P.graphs) (void)hipGraphExecDestroy(kv.second);
    P.graphs.clear();
  }
}
}  // namespace

extern "C" int wg_load_forcing(wg_handle *h, const float *forcing, int32_t n_days) {
  if (!h || !forcing || n_days <= 0) return WG_E_ARG;
  const int32_t n = h->n;
  const size_t day_elems = size_t(4) * n;
  // the reference reads the forcing grids as given; values that no climate input
  // holds (not finite, negative precipitation) are rejected before anything changes
  for (int32_t d = 0; d < n_days; d++) {
    const float *src = forcing + day_elems * d;
    for (size_t k = 0; k < day_elems; k++)
      if (!std::isfinite(src[k]))
        return fail(h, WG_E_INPUT, "forcing day %d, variable %d, cell %d is not finite", d, int(k / n), int(k % n));
    for (int32_t i = 0; i < n; i++)
      if (src[size_t(n) + i] < 0.f)
        return fail(h, WG_E_INPUT, "forcing day %d: negative precipitation %g at cell %d", d, src[size_t(n) + i], i);
  }
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  // the captured graphs hold the old buffer's address
  clear_graphs(h);
  float *fresh = nullptr;
  hipError_t e = hipMalloc(&fresh, day_elems * n_days * sizeof(float));
  if (e != hipSuccess) return fail(h, WG_E_NOMEM, "forcing buffer of %d days: %s", n_days, hipGetErrorString(e));
  std::vector<float> buf(day_elems);
  for (int32_t d = 0; d < n_days && e == hipSuccess; d++) {
    const float *src = forcing + day_elems * d;
    for (int v = 0; v < 4; v++)
      for (int32_t p = 0; p < n; p++) buf[size_t(v) * n + p] = src[size_t(v) * n + h->perm[p]];
    e = hipMemcpy(fresh + day_elems * d, buf.data(), day_elems * sizeof(float), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    (void)hipFree(fresh);
    return fail(h, WG_E_HIP, "forcing upload: %s", hipGetErrorString(e));
  }
  if (h->d_forcing) (void)hipFree(h->d_forcing);
  h->d_forcing = fresh;
  h->n_forcing_days = n_days;
  return WG_OK;
}

extern "C" int wg_set_stations(wg_handle *h, int32_t n_st, const int32_t *cells) {
  if (!h || n_st < 0 || (n_st > 0 && !cells)) return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  std::vector<int32_t> slot(h->n, -1);
  for (int32_t b = 0; b < n_st; b++) {
    if (cells[b] < 0 || cells[b] >= h->n) return fail(h, WG_E_ARG, "station cell %d out of range", cells[b]);
    // the reference stops at the first station of a cell (routing.cpp:4232-4238)
    if (slot[h->inv[cells[b]]] < 0) slot[h->inv[cells[b]]] = b;
  }
  // the new accumulators first: on failure the handle keeps its stations
  auto release = [&](void *ptr) { release_buf(h, ptr); };
  double *st_sum = nullptr, *st_daily = nullptr;
  int rc = dev_alloc(h, &st_sum, size_t(h->m) * std::max(n_st, 1));
  if (!rc) rc = dev_alloc(h, &st_daily, size_t(h->m) * WG_DAILY_SLOTS * std::max(n_st, 1));
  if (rc) {
    if (st_sum) release(st_sum);
    return rc;
  }
  // the captured graphs hold the accumulators' addresses and the station count
  clear_graphs(h);
  for (int32_t p = 0; p < h->n; p++) h->clink[p].station = slot[p];
  const std::vector<WgI4> packed = pack_links(h->clink);
  hipError_t e = hipMemcpy(const_cast<int32_t *>(h->S.station), slot.data(), h->n * sizeof(int32_t),
                           hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(const_cast<WgI4 *>(h->S.clink), packed.data(), packed.size() * sizeof(WgI4),
                  hipMemcpyHostToDevice);
  release(h->X.st_sum);
  release(h->X.st_daily);
  h->X.st_sum = st_sum;
  h->X.st_daily = st_daily;
  h->n_stations = n_st;
  h->S.n_stations = n_st;
  h->station_cells.assign(cells, cells + n_st);
  h->last_daily_days = 0;
  if (e != hipSuccess) return fail(h, WG_E_HIP, "station tables: %s", hipGetErrorString(e));
  return WG_OK;
}

namespace {

// Launches the cell kernel with the handle's snow chunk.
void launch_cell(wg_handle *h, const WgStatic &S, hipStream_t stream, dim3 grid, int k, int do_post,
                 int do_vertical, int write_diag, int slot) {
  // the kernels without water use carry no use code
  auto *kern = h->snow_chunk == 10 ? (S.use_on ? wg_cell_kernel<10, true> : wg_cell_kernel<10, false>)
                                   : (S.use_on ? wg_cell_kernel<5, true> : wg_cell_kernel<5, false>);
  hipLaunchKernelGGL(kern, grid, dim3(h->block_threads_cell), 0, stream, S, h->X, h->T, h->d_forcing,
                     h->n_forcing_days, h->d_clock, k, do_post, do_vertical, write_diag, slot);
}

// Enqueues one block of `days` days (captured or eager): of the whole grid on the
// handle's stream (part -1), or of one partition on its stream, without the
// station sums, which the caller folds once every partition is done.
int enqueue_block(wg_handle *h, int days, hipEvent_t *evs, int part = -1) {
  const int32_t M = h->m;
  WgStatic S = h->S;
  hipStream_t stream = h->stream;
  uint32_t n_tiles = uint32_t(h->tile_start.size()), n_blocks = uint32_t(h->block_group.size());
  int32_t c0 = 0, c1 = h->n;
  if (part >= 0) {
    const WgPart &P = h->parts[part];
    S.tile_start += P.tile_off;
    S.tile_end += P.tile_off;
    S.block_group += P.group_block_off;
    stream = P.stream;
    n_tiles = uint32_t(P.n_tiles);
    n_blocks = uint32_t(P.n_group_blocks);
    c0 = P.cell_begin;
    c1 = P.cell_end;
  }
  const dim3 cgrid(n_tiles, M);
  const dim3 chgrid(n_blocks, M);
  int ev = 0;
  for (int k = 0; k < days; k++) {
    const int slot = k % WG_DAILY_SLOTS;
    if (evs) HIP_CHECK(h, hipEventRecord(evs[ev++], stream));
    launch_cell(h, S, stream, cgrid, k, k > 0 ? 1 : 0, 1, k == days - 1 ? 1 : 0, slot);
    HIP_CHECK(h, hipGetLastError());
    if (evs) HIP_CHECK(h, hipEventRecord(evs[ev++], stream));
    if (h->chain_mode == 3) {
      // the wave walk runs 8 wavefronts with a 512-thread configuration, else 4
      const bool w8 = h->block_threads_chain >= 512;
      auto *kern = w8 ? (S.use_on ? wg_chain_wave_kernel<8, true> : wg_chain_wave_kernel<8, false>)
                      : (S.use_on ? wg_chain_wave_kernel<4, true> : wg_chain_wave_kernel<4, false>);
      hipLaunchKernelGGL(kern, chgrid, dim3(w8 ? 512 : 256), wave_lds_bytes(h), stream, S, h->X, h->T,
                         h->d_level_off, h->d_group_first, h->d_wave_wide, h->d_wave_item_off[w8 ? 1 : 0],
                         h->d_wave_items[w8 ? 1 : 0], h->d_clock, k, slot);
    }
    else if (h->chain_mode == 1)
      hipLaunchKernelGGL(S.use_on ? wg_chain_flow_kernel<true> : wg_chain_flow_kernel<false>, chgrid,
                         dim3(h->block_threads_chain), size_t(h->max_group_cells) * sizeof(double), stream, S,
                         h->X, h->T, h->d_level_off, h->d_group_first, h->d_clock, k, slot);
    else {
      auto *kern = h->chain_mode == 2 ? (S.use_on ? wg_chain_kernel<2, true> : wg_chain_kernel<2, false>)
                                      : (S.use_on ? wg_chain_kernel<1, true> : wg_chain_kernel<1, false>);
      hipLaunchKernelGGL(kern, chgrid, dim3(h->block_threads_chain), size_t(h->max_group_cells) * sizeof(double),
                         stream, S, h->X, h->T, h->d_level_off, h->d_group_first, h->d_clock, k, slot);
    }
    HIP_CHECK(h, hipGetLastError());
    if (h->X.pot_acc) {
      hipLaunchKernelGGL(wg_pot_runoff_kernel, dim3((c1 - c0 + 255) / 256, M), dim3(256), 0, stream, S, h->X, c0,
                         c1);
      HIP_CHECK(h, hipGetLastError());
    }
    if (part < 0 && h->n_stations > 0 && (slot == WG_DAILY_SLOTS - 1 || k == days - 1)) {
      // fold the daily station series into the running sums, in day order
      const int nsm = h->n_stations * M;
      hipLaunchKernelGGL(wg_station_reduce, dim3((nsm + 255) / 256), dim3(256), 0, stream, h->X,
                         h->n_stations, M, 0, slot + 1);
      HIP_CHECK(h, hipGetLastError());
    }
  }
  if (evs) HIP_CHECK(h, hipEventRecord(evs[ev++], stream));
  // closing pass: finish the last day (day `days` - 1) so that the state is complete
  launch_cell(h, S, stream, cgrid, days, 1, 0, 0, 0);
  HIP_CHECK(h, hipGetLastError());
  if (evs) HIP_CHECK(h, hipEventRecord(evs[ev++], stream));
  return WG_OK;
}

int get_graph(wg_handle *h, int days, hipGraphExec_t *out, int part = -1) {
  std::map<int, hipGraphExec_t> &graphs = part < 0 ? h->graphs : h->parts[part].graphs;
  hipStream_t stream = part < 0 ? h->stream : h->parts[part].stream;
  auto it = graphs.find(days);
  if (it != graphs.end()) {
    *out = it->second;
    return WG_OK;
  }
  hipGraph_t graph;
  HIP_CHECK(h, hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_block(h, days, nullptr, part);
  hipError_t e = hipStreamEndCapture(stream, &graph);
  if (rc) return rc;
  if (e != hipSuccess) return fail(h, WG_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  hipGraphExec_t exec;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(h, WG_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  graphs[days] = exec;
  *out = exec;
  return WG_OK;
}

int check_step_args(wg_handle *h, int32_t forcing_day0, int32_t first_doy, int32_t n_days) {
  if (n_days <= 0 || first_doy < 1 || first_doy > 365 || forcing_day0 < 0)
    return fail(h, WG_E_ARG, "invalid day range");
  if (!h->d_forcing) return fail(h, WG_E_STATE, "no forcing loaded");
  if (int64_t(forcing_day0) + n_days > h->n_forcing_days)
    return fail(h, WG_E_ARG, "forcing days %d..%d exceed the %lld loaded days", forcing_day0,
                forcing_day0 + n_days - 1, (long long)h->n_forcing_days);
  if (h->subtract_use != 0) {
    if (!h->S.use_on) return fail(h, WG_E_STATE, "water use is on but no tables are set (wg_set_water_use)");
    if (h->use_gen != h->statics_gen)
      return fail(h, WG_E_STATE, "wg_annual_init changed the statics: set the water-use tables again");
    const int64_t first = int64_t(forcing_day0) - h->S.use_day0, last = first + n_days - 1;
    if (first < 0 || last >= int64_t(h->S.use_years) * 365)
      return fail(h, WG_E_ARG, "days %d..%d are outside the water-use tables", forcing_day0, forcing_day0 + n_days - 1);
  }
  return WG_OK;
}

// The dataflow chain kernel flags a walk that hit its spin limit (it cannot for a
// valid schedule); the day's routing is then incomplete and the call fails.
int check_flow_error(wg_handle *h) {
  if (h->chain_mode != 1) return WG_OK;
  int32_t flag = 0;
  HIP_CHECK(h, hipMemcpy(&flag, h->X.flow_error, sizeof(int32_t), hipMemcpyDeviceToHost));
  if (flag) {
    const int32_t zero = 0;
    HIP_CHECK(h, hipMemcpy(h->X.flow_error, &zero, sizeof(int32_t), hipMemcpyHostToDevice));
    return fail(h, WG_E_STATE, "dataflow routing walk exceeded its spin limit");
  }
  return WG_OK;
}

}  // namespace

extern "C" int wg_step_days(wg_handle *h, int32_t forcing_day0, int32_t first_doy, int32_t n_days) {
  if (!h) return WG_E_ARG;
  int rc = check_step_args(h, forcing_day0, first_doy, n_days);
  if (rc) return rc;
  HIP_CHECK(h, hipSetDevice(h->device));
  // a block's daily station series fills the WG_DAILY_SLOTS slots in day order
  int block = 365;
  if (const char *v = getenv("WG_GRAPH_DAYS")) block = std::max(1, std::min(WG_DAILY_SLOTS, atoi(v)));
  HIP_CHECK(h, hipEventRecord(h->ev0, h->stream));
  int64_t day = forcing_day0;
  int doy = first_doy;
  int remaining = n_days;
  int64_t launches = 0;
  const int n_parts = int(h->parts.size());
  while (remaining > 0) {
    const int k = std::min(block, remaining);
    hipLaunchKernelGGL(wg_clock_set, dim3(1), dim3(64), 0, h->stream, h->d_clock, day, doy, h->started ? 1 : 0);
    HIP_CHECK(h, hipGetLastError());
    if (n_parts < 2) {
      hipGraphExec_t g;
      rc = get_graph(h, k, &g);
      if (rc) return rc;
      HIP_CHECK(h, hipGraphLaunch(g, h->stream));
    } else {
      // every partition's block on its own stream after the clock, then the
      // station sums once all are done (the sums of a block are folded in day order)
      for (int pp = 0; pp < n_parts; pp++) {
        hipGraphExec_t g;
        rc = get_graph(h, k, &g, pp);
        if (rc) return rc;
      }
      HIP_CHECK(h, hipEventRecord(h->ev_fork, h->stream));
      for (WgPart &P : h->parts) {
        HIP_CHECK(h, hipStreamWaitEvent(P.stream, h->ev_fork, 0));
        HIP_CHECK(h, hipGraphLaunch(P.graphs[k], P.stream));
        HIP_CHECK(h, hipEventRecord(P.done, P.stream));
        HIP_CHECK(h, hipStreamWaitEvent(h->stream, P.done, 0));
      }
      if (h->n_stations > 0) {
        const int nsm = h->n_stations * h->m;
        hipLaunchKernelGGL(wg_station_reduce, dim3((nsm + 255) / 256), dim3(256), 0, h->stream, h->X,
                           h->n_stations, h->m, 0, (k - 1) % WG_DAILY_SLOTS + 1);
        HIP_CHECK(h, hipGetLastError());
      }
    }
    h->started = true;
    day += k;
    doy = (doy - 1 + k) % 365 + 1;
    remaining -= k;
    launches += 2 * k + 1;
  }
  HIP_CHECK(h, hipEventRecord(h->ev1, h->stream));
  HIP_CHECK(h, hipEventSynchronize(h->ev1));
  float ms = 0;
  HIP_CHECK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms[2] = ms;
  h->last_launches[2] = launches;
  h->last_daily_days = (n_days - 1) % block + 1;
  h->last_daily_first = n_days - h->last_daily_days;
  return check_flow_error(h);
}

// Eager variant with an event pair around every kernel: per-kernel durations for
// the roofline figures (wg_last_timing).
extern "C" int wg_step_days_profiled(wg_handle *h, int32_t forcing_day0, int32_t first_doy, int32_t n_days) {
  if (!h) return WG_E_ARG;
  int rc = check_step_args(h, forcing_day0, first_doy, n_days);
  if (rc) return rc;
  if (n_days > WG_DAILY_SLOTS)
    return fail(h, WG_E_ARG, "the profiled path runs at most %d days per call", WG_DAILY_SLOTS);
  HIP_CHECK(h, hipSetDevice(h->device));
  std::vector<hipEvent_t> evs(size_t(2) * n_days + 2);
  for (auto &e : evs) HIP_CHECK(h, hipEventCreate(&e));
  hipLaunchKernelGGL(wg_clock_set, dim3(1), dim3(64), 0, h->stream, h->d_clock, int64_t(forcing_day0), first_doy,
                     h->started ? 1 : 0);
  rc = enqueue_block(h, n_days, evs.data());
  if (rc) return rc;
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  h->started = true;
  double cell = 0, chain = 0;
  for (int32_t k = 0; k < n_days; k++) {
    float a = 0, b = 0;
    HIP_CHECK(h, hipEventElapsedTime(&a, evs[2 * k], evs[2 * k + 1]));
    HIP_CHECK(h, hipEventElapsedTime(&b, evs[2 * k + 1], evs[2 * k + 2]));
    cell += a;
    chain += b;
  }
  float total = 0;
  HIP_CHECK(h, hipEventElapsedTime(&total, evs[0], evs.back()));
  for (auto &e : evs) (void)hipEventDestroy(e);
  h->last_ms[0] = cell;
  h->last_ms[1] = chain;
  h->last_ms[2] = total;
  h->last_launches[0] = n_days;
  h->last_launches[1] = n_days;
  h->last_launches[2] = 2 * int64_t(n_days) + 1;
  h->last_daily_days = n_days;
  h->last_daily_first = 0;
  return check_flow_error(h);
}

namespace {

int apply_config(wg_handle *h, int cell_threads, int chain_threads, int chunk, int chain_mode) {
  if (chain_mode == 1 && h->max_depth > WG_MAX_GROUP_DEPTH)
    return fail(h, WG_E_ARG, "dataflow chain kernel supports groups of up to %d levels", WG_MAX_GROUP_DEPTH);
  h->block_threads_cell = cell_threads;
  h->block_threads_chain = chain_threads;
  h->snow_chunk = chunk;
  h->chain_mode = chain_mode;
  build_tiles(h);
  clear_graphs(h);
  return upload_tiles(h);
}

}  // namespace

// Tries cell-kernel tiles of 128 and 256 cells, the chain modes with their
// workgroups (level walks 64, 128 or 256 threads, wave walk 4 or 8 wavefronts)
// and snow chunks of 5 and 10 sub-cell pairs on n_days days of the loaded
// forcing, each from the same saved state and run as wg_step_days runs them, and
// keeps the fastest.  Results do not depend on the choice (every configuration is
// bit-identical).
extern "C" int wg_autotune(wg_handle *h, int32_t n_days, int32_t *config, double *ms_per_day) {
  if (!h || n_days <= 0) return WG_E_ARG;
  if (n_days > h->n_forcing_days) return fail(h, WG_E_ARG, "autotune needs %d loaded forcing days", n_days);
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  size_t total = 0;
  for (const auto &b : h->bufs) total += b.bytes;
  char *snap = nullptr;
  if (hipMalloc(&snap, total) != hipSuccess) return fail(h, WG_E_NOMEM, "autotune snapshot of %zu bytes", total);
  auto copy_all = [&](bool save) -> hipError_t {
    size_t off = 0;
    for (const auto &b : h->bufs) {
      hipError_t e = save ? hipMemcpyAsync(snap + off, b.p, b.bytes, hipMemcpyDeviceToDevice, h->stream)
                          : hipMemcpyAsync(b.p, snap + off, b.bytes, hipMemcpyDeviceToDevice, h->stream);
      if (e != hipSuccess) return e;
      off += b.bytes;
    }
    return hipStreamSynchronize(h->stream);
  };
  const bool started = h->started;
  const int32_t daily = h->last_daily_days;
  const int c0 = h->block_threads_cell, ch0 = h->block_threads_chain, k0 = h->snow_chunk, m0 = h->chain_mode;
  int rc = WG_OK;
  if (copy_all(true) != hipSuccess) {
    (void)hipFree(snap);
    return fail(h, WG_E_HIP, "autotune snapshot copy");
  }
  const int cells[] = {128, 256}, chains[] = {64, 128, 256, 512}, chunks[] = {5, 10}, modes[] = {0, 1, 2, 3};
  double best = 1e300;
  int bc = h->block_threads_cell, bch = h->block_threads_chain, bk = h->snow_chunk, bm = h->chain_mode;
  for (int c : cells)
    for (int ch : chains)
      for (int k : chunks)
        for (int md : modes) {
        if (rc) break;
        if (md == 1 && h->max_depth > WG_MAX_GROUP_DEPTH) continue;
        // the wave walk runs 4 or 8 wavefronts (256 or 512 threads); the level walks up to 256 threads
        if (md == 3 ? (ch != 256 && ch != 512) : ch > 256) continue;
        rc = apply_config(h, c, ch, k, md);
        double ms = 0;
        // the first pass loads the kernels and captures the graphs of the path that
        // wg_step_days runs (with its partitions); the second is timed
        for (int rep = 0; rep < 2 && !rc; rep++) {
          if (copy_all(false) != hipSuccess) rc = fail(h, WG_E_HIP, "autotune restore");
          if (!rc) rc = upload_tiles(h);
          h->started = started;
          if (!rc) rc = wg_step_days(h, 0, 1, n_days);
          ms = h->last_ms[2];
        }
        if (!rc && ms < best) {
          best = ms;
          bc = c;
          bch = ch;
          bk = k;
          bm = md;
        }
      }
  // the state and (after a failure) the configuration are restored on every path
  const std::string err = h->err;
  int rc2 = rc ? apply_config(h, c0, ch0, k0, m0) : apply_config(h, bc, bch, bk, bm);
  if (copy_all(false) != hipSuccess && !rc2) rc2 = fail(h, WG_E_HIP, "autotune restore");
  if (!rc2) rc2 = upload_tiles(h);
  h->started = started;
  h->last_daily_days = daily;
  (void)hipFree(snap);
  if (rc) {
    h->err = err;
    return rc;
  }
  if (rc2) return rc2;
  if (config) {
    config[0] = bc;
    config[1] = bch;
    config[2] = bk;
    config[3] = bm;
  }
  if (ms_per_day) *ms_per_day = best / n_days;
  return WG_OK;
}

// Sets the launch configuration directly (cell tile 64-256 in steps of 64, chain
// workgroup 64-256, or up to 512 for the wave walk, snow chunk 5 or 10).
extern "C" int wg_set_launch_config(wg_handle *h, int32_t cell_threads, int32_t chain_threads, int32_t snow_chunk,
                                    int32_t chain_mode) {
  if (!h || cell_threads < 64 || cell_threads > 256 || cell_threads % 64 || chain_threads < 64 ||
      chain_threads > (chain_mode == 3 ? 512 : 256) || (snow_chunk != 5 && snow_chunk != 10) || chain_mode < 0 ||
      chain_mode > 3)
    return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return apply_config(h, cell_threads, chain_threads, snow_chunk, chain_mode);
}

extern "C" int wg_get_station_discharge(wg_handle *h, double *sum, double *daily, int32_t max_days) {
  if (!h) return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  const int32_t ns = h->n_stations;
  if (ns == 0) return WG_OK;
  if (sum) HIP_CHECK(h, hipMemcpy(sum, h->X.st_sum, size_t(h->m) * ns * sizeof(double), hipMemcpyDeviceToHost));
  if (daily) {
    const int32_t nd = std::min(max_days, h->last_daily_days);
    std::vector<double> buf(size_t(h->m) * WG_DAILY_SLOTS * ns);
    HIP_CHECK(h, hipMemcpy(buf.data(), h->X.st_daily, buf.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int32_t mm = 0; mm < h->m; mm++)
      for (int32_t d = 0; d < nd; d++)
        for (int32_t b = 0; b < ns; b++)
          daily[(size_t(mm) * max_days + d) * ns + b] = buf[(size_t(mm) * WG_DAILY_SLOTS + d) * ns + b];
  }
  return WG_OK;
}

extern "C" int wg_station_daily_window(wg_handle *h, int32_t *first_day, int32_t *n_days) {
  if (!h || !first_day || !n_days) return WG_E_ARG;
  *first_day = h->last_daily_first;
  *n_days = h->last_daily_days;
  return WG_OK;
}

namespace {

// Chain flags of reference cell i (the links' flags, with the reservoir type and
// start month in bits 8-23 when with_res_bits).
int chain_flags(const wg_static_fields *st, int32_t i, bool with_res_bits) {
  int f = 0;
  if (st->lake_area[i] > 0.) f |= CFL_LAKE;
  if (st->reservoir_area[i] > 0.) f |= CFL_RES;
  if (st->glo_wetland[i] > 0) f |= CFL_GWET;
  if (st->aindex[i] == 1 && st->ldd[i] >= 0) f |= CFL_ARID;
  if (st->ldd[i] >= 0) f |= CFL_LDD;
  if (!(st->contfreq[i] < 0.000001)) f |= CFL_CONT;
  if (with_res_bits && st->reservoir_area[i] > 0.)
    f |= (int(st->res_type[i]) & 0xff) << 8 | (int(st->start_month[i]) & 0xff) << 16;
  return f;
}

template <class T>
int to_host(wg_handle *h, std::vector<T> &out, const T *dev, size_t count) {
  out.resize(count);
  HIP_CHECK(h, hipMemcpy(out.data(), dev, count * sizeof(T), hipMemcpyDeviceToHost));
  return WG_OK;
}

template <class T>
int to_device(wg_handle *h, const T *dev, const std::vector<T> &in) {
  HIP_CHECK(h, hipMemcpy(const_cast<T *>(dev), in.data(), in.size() * sizeof(T), hipMemcpyHostToDevice));
  return WG_OK;
}

}  // namespace

extern "C" int wg_annual_init(wg_handle *h, const wg_static_fields *st, const int32_t *outflow_cell,
                              int32_t apply_changes) {
  if (!h || !st) return WG_E_ARG;
  const int32_t n = h->n, M = h->m;
  const void *req[] = {st->cell_area, st->contfreq, st->aindex, st->ldd, st->glo_wetland, st->lake_area,
                       st->reservoir_area, st->glo_res, st->stor_cap, st->mean_outflow, st->res_type,
                       st->start_month, st->downstream};
  for (const void *q : req)
    if (!q) return fail(h, WG_E_ARG, "wg_annual_init: a required static field is NULL");
  for (int32_t i = 0; i < n; i++) {
    if (st->reservoir_area[i] > 0. && !(st->res_type[i] == 1 || st->res_type[i] == 2))
      return fail(h, WG_E_INPUT, "cell %d: reservoir of unknown type (routing.cpp:1106-1115)", i);
    if (st->reservoir_area[i] > 0. && (st->start_month[i] < 1 || st->start_month[i] > 12))
      return fail(h, WG_E_INPUT, "cell %d: reservoir start month out of range", i);
  }
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  // Part 4 of routingClass::annualInit (routing.cpp:1300-1421): a cell whose
  // reservoir fraction grows gives the new reservoir its land: the land-area
  // fraction falls by the growth, and the water of the lost land (canopy, soil and
  // snow, mm over the land area) moves to the reservoir storage of the cell's
  // outflow cell assignment.  Cells in reference order, every member.
  std::vector<int32_t> grown;
  if (apply_changes) {
    for (int32_t i = 0; i < n; i++)
      if (st->glo_res[i] - h->glo_res_prev[i] > 0.) grown.push_back(i);
  }
  if (!grown.empty()) {
    if (!outflow_cell) return fail(h, WG_E_ARG, "wg_annual_init: outflow cell assignment missing");
    for (int32_t i : grown)
      if (outflow_cell[i] < 1 || outflow_cell[i] > n)
        return fail(h, WG_E_INPUT, "cell %d: outflow cell assignment %d out of range", i, outflow_cell[i]);
    const size_t NM = size_t(n) * M;
    std::vector<double> canopy, soil, snow, laf, laf_prev, res;
    int rc = to_host(h, canopy, h->X.canopy, NM);
    if (!rc) rc = to_host(h, soil, h->X.soil, NM);
    if (!rc) rc = to_host(h, snow, h->X.snow, NM);
    if (!rc) rc = to_host(h, laf, h->X.laf, NM);
    if (!rc) rc = to_host(h, laf_prev, h->X.laf_prev, NM);
    if (!rc) rc = to_host(h, res, h->X.res, NM);
    if (rc) return rc;
    // every member's new fractions first: a cell whose land did not shrink against
    // the day before takes a branch (routing.cpp:1349-1357) that keeps the next
    // day's fraction apart from today's, which this state does not hold
    for (int32_t m = 0; m < M; m++) {
      for (int32_t i : grown) {
        const size_t mp = size_t(m) * n + h->inv[i];
        const double change = st->glo_res[i] - h->glo_res_prev[i];
        double f = h->glo_res_prev[i] > 0. ? laf[mp] - change : laf[mp] - st->glo_res[i];
        if (f < 0.) f = 0.;
        if (f - laf_prev[mp] >= 0.)
          return fail(h, WG_E_INPUT, "cell %d member %d: the land fraction did not fall against the day before "
                      "(routing.cpp:1354-1357 is not supported)", i, m);
      }
    }
    for (int32_t m = 0; m < M; m++) {
      for (int32_t i : grown) {
        const size_t mp = size_t(m) * n + h->inv[i];
        const size_t ma = size_t(m) * n + h->inv[outflow_cell[i] - 1];
        const double change = st->glo_res[i] - h->glo_res_prev[i];
        double f = h->glo_res_prev[i] > 0. ? laf[mp] - change : laf[mp] - st->glo_res[i];
        if (f < 0.) f = 0.;
        const double lost = f - laf_prev[mp];
        const double area = st->cell_area[i];
        const double c = canopy[mp] * area / 1000000.0 * ((-lost) / 100.0);
        const double so = soil[mp] * area / 1000000.0 * ((-lost) / 100.0);
        const double sn = snow[mp] * area / 1000000.0 * ((-lost) / 100.0);
        res[ma] += c + so + sn;
        laf[mp] = f;
      }
    }
    rc = to_device(h, h->X.laf, laf);
    if (!rc) rc = to_device(h, h->X.res, res);
    if (rc) return rc;
  }
  // Parts 1-2 (routing.cpp:1015-1100): this year's reservoirs, as the caller derived
  // them from the full grids and the start years; flags, tables and the release
  // terms follow them.  The device order stays (cells with a reservoir last in
  // their level is a speed choice, not a condition).
  int rc = to_device(h, h->S.reservoir_area, permute(h, st->reservoir_area));
  if (!rc) rc = to_device(h, h->S.stor_cap, permute(h, st->stor_cap));
  if (!rc) rc = to_device(h, h->S.res_prov_rel, permute(h, reservoir_prov_rel(st, n).data()));
  if (rc) return rc;
  h->statics_gen++;
  std::vector<int8_t> cflags(n);
  for (int32_t p = 0; p < n; p++) {
    const int32_t i = h->perm[p];
    cflags[p] = int8_t(chain_flags(st, i, false));
    h->clink[p].flags = chain_flags(st, i, true);
  }
  rc = to_device(h, h->S.cflags, cflags);
  if (!rc) rc = to_device(h, h->S.clink, pack_links(h->clink));
  for (int t = 0; t < 2 && !rc; t++) {
    std::vector<int32_t> item_off;
    std::vector<WgWaveItem> items;
    wave_tables(h, t == 0 ? 4 : 8, item_off, items);
    rc = to_device(h, h->d_wave_items[t], items);
  }
  if (rc) return rc;
  // the potential cell runoff restarts with the year (routing.cpp:1009)
  if (h->X.pot_acc) {
    HIP_CHECK(h, hipMemsetAsync(h->X.pot_acc, 0, size_t(h->m) * n * sizeof(double), h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
  }
  h->glo_res_prev.assign(st->glo_res, st->glo_res + n);
  return WG_OK;
}

extern "C" int wg_set_pot_cell_runoff(wg_handle *h, int32_t enable) {
  if (!h) return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  const size_t mn = size_t(h->m) * h->n;
  if (enable && h->X.pot_acc) {
    HIP_CHECK(h, hipMemsetAsync(h->X.pot_acc, 0, mn * sizeof(double), h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return WG_OK;
  }
  if (enable) {
    double *buf[4] = {nullptr, nullptr, nullptr, nullptr};
    int rc = WG_OK;
    for (int i = 0; i < 4 && !rc; i++) rc = dev_alloc(h, &buf[i], mn);
    if (rc) {
      for (double *b : buf)
        if (b) release_buf(h, b);
      return rc;
    }
    // the captured days hold the state's pointers
    clear_graphs(h);
    h->X.pot_acc = buf[0];
    h->X.pot_base = buf[1];
    h->X.pot_ow = buf[2];
    h->X.pot_lgw = buf[3];
  } else if (h->X.pot_acc) {
    clear_graphs(h);
    for (double *b : {h->X.pot_acc, h->X.pot_base, h->X.pot_ow, h->X.pot_lgw}) release_buf(h, b);
    h->X.pot_acc = h->X.pot_base = h->X.pot_ow = h->X.pot_lgw = nullptr;
  }
  return WG_OK;
}

extern "C" int wg_set_water_use(wg_handle *h, const wg_static_fields *st, int32_t n_years, int32_t year0_day,
                                const double *nus_m3, const double *nug_m3, const double *withdrawal_irrig_m3,
                                const double *consumptive_irrig_m3, const double *frac_return_gw, int32_t delayed) {
  if (!h || !st || n_years <= 0 || year0_day < -365 * 1000 || !nus_m3 || !nug_m3 || !withdrawal_irrig_m3 ||
      !consumptive_irrig_m3 || !frac_return_gw || (delayed != 0 && delayed != 1) || !st->reservoir_area ||
      !st->res_type || !st->mean_outflow || !st->downstream || !st->params)
    return WG_E_ARG;
  if (h->subtract_use == 0) return fail(h, WG_E_STATE, "the handle was made with subtract_use 0");
  const int32_t n = h->n, M = h->m;
  const size_t tab = size_t(n_years) * 12 * n;
  static const int kDays[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  // dailyNUInit (routing.cpp:909-972): NUs and NUg times their multipliers, then
  // over 1e9 x days of the month; the irrigation withdrawal and consumptive use
  // over the same (routing.cpp:3918-3919).  Reference order [y][month][n] first.
  std::vector<double> nus(tab), nug(tab), wu(tab), cu(tab), prov(tab);
  const double *mult_s = st->params + size_t(WG_M_NETABSSW) * n, *mult_g = st->params + size_t(WG_M_NETABSGW) * n;
  for (int32_t y = 0; y < n_years; y++)
    for (int mo = 0; mo < 12; mo++) {
      const double div = 1000000000. * double(kDays[mo]);
      double *a = &nus[(size_t(y) * 12 + mo) * n], *b = &nug[(size_t(y) * 12 + mo) * n];
      double *c = &wu[(size_t(y) * 12 + mo) * n], *d = &cu[(size_t(y) * 12 + mo) * n];
      for (int32_t i = 0; i < n; i++) {
        const size_t k = (size_t(y) * n + i) * 12 + mo;
        const double vs = mult_s[i] * nus_m3[k], vg = mult_g[i] * nug_m3[k];
        a[i] = vs / div;
        b[i] = vg / div;
        c[i] = withdrawal_irrig_m3[k] / div;
        d[i] = consumptive_irrig_m3[k] / div;
      }
      const std::vector<double> pr = reservoir_prov_rel(st, n, a, kDays[mo]);
      std::copy(pr.begin(), pr.end(), &prov[(size_t(y) * 12 + mo) * n]);
    }
  // to the device order, table by table
  auto dev_order = [&](std::vector<double> &v) {
    std::vector<double> t(static_cast<size_t>(n));
    for (size_t r = 0; r < size_t(n_years) * 12; r++) {
      double *row = &v[r * n];
      for (int32_t p = 0; p < n; p++) t[p] = row[h->perm[p]];
      std::copy(t.begin(), t.end(), row);
    }
  };
  for (std::vector<double> *v : {&nus, &nug, &wu, &cu, &prov}) dev_order(*v);
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  double *d_tab[5] = {nullptr, nullptr, nullptr, nullptr, nullptr}, *d_frgi = nullptr, *d_use = nullptr;
  const bool fresh = h->X.use == nullptr;
  int rc = WG_OK;
  for (int i = 0; i < 5 && !rc; i++) rc = dev_alloc(h, &d_tab[i], tab);
  if (!rc) rc = dev_alloc(h, &d_frgi, size_t(n));
  if (!rc && fresh) rc = dev_alloc(h, &d_use, size_t(WG_USE_NFIELDS + 1) * M * n);
  if (rc) {
    for (double *b : d_tab)
      if (b) release_buf(h, b);
    if (d_frgi) release_buf(h, d_frgi);
    return rc;
  }
  const std::vector<double> *src[5] = {&nus, &nug, &wu, &cu, &prov};
  hipError_t e = hipSuccess;
  for (int i = 0; i < 5 && e == hipSuccess; i++)
    e = hipMemcpy(d_tab[i], src[i]->data(), tab * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const std::vector<double> fr = permute(h, frac_return_gw);
    e = hipMemcpy(d_frgi, fr.data(), size_t(n) * sizeof(double), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    for (double *b : d_tab) release_buf(h, b);
    release_buf(h, d_frgi);
    if (d_use) release_buf(h, d_use);
    return fail(h, WG_E_HIP, "water-use tables: %s", hipGetErrorString(e));
  }
  // the captured days hold the statics' and the state's pointers
  clear_graphs(h);
  for (const double *b : {h->S.use_nus, h->S.use_nug, h->S.use_wu, h->S.use_cu, h->S.use_provrel, h->S.use_frgi})
    if (b) release_buf(h, const_cast<double *>(b));
  h->S.use_nus = d_tab[0];
  h->S.use_nug = d_tab[1];
  h->S.use_wu = d_tab[2];
  h->S.use_cu = d_tab[3];
  h->S.use_provrel = d_tab[4];
  h->S.use_frgi = d_frgi;
  if (fresh) {
    h->X.use = d_use;
    h->X.use_rem = d_use + size_t(WG_USE_NFIELDS) * M * n;
  }
  h->S.use_stride = int64_t(M) * n;
  h->S.use_years = n_years;
  h->S.use_day0 = year0_day;
  h->S.use_delayed = delayed;
  h->S.use_on = 1;
  h->use_gen = h->statics_gen;
  return WG_OK;
}

extern "C" int wg_get_water_use_state(wg_handle *h, int32_t member, double *state) {
  if (!h || !state || member < 0 || member >= h->m) return WG_E_ARG;
  if (!h->X.use) return fail(h, WG_E_STATE, "water use is not on");
  const int32_t n = h->n;
  std::vector<double> t(size_t(WG_USE_NFIELDS) * n);
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  for (int f = 0; f < WG_USE_NFIELDS; f++)
    HIP_CHECK(h, hipMemcpy(&t[size_t(f) * n], h->X.use + (size_t(f) * h->m + member) * n, size_t(n) * sizeof(double),
                           hipMemcpyDeviceToHost));
  for (int f = 0; f < WG_USE_NFIELDS; f++)
    for (int32_t p = 0; p < n; p++) state[size_t(f) * n + h->perm[p]] = t[size_t(f) * n + p];
  return WG_OK;
}

extern "C" int wg_set_water_use_state(wg_handle *h, int32_t member, const double *state) {
  if (!h || !state || member < 0 || member >= h->m) return WG_E_ARG;
  if (!h->X.use) return fail(h, WG_E_STATE, "water use is not on");
  const int32_t n = h->n;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  for (int f = 0; f < WG_USE_NFIELDS; f++) {
    const std::vector<double> v = permute(h, state + size_t(f) * n);
    HIP_CHECK(h, hipMemcpy(h->X.use + (size_t(f) * h->m + member) * n, v.data(), size_t(n) * sizeof(double),
                           hipMemcpyHostToDevice));
  }
  return WG_OK;
}

extern "C" int wg_get_pot_cell_runoff(wg_handle *h, int32_t member, double *pot) {
  if (!h || !pot || member < 0 || member >= h->m) return WG_E_ARG;
  if (!h->X.pot_acc) return fail(h, WG_E_ARG, "the potential cell runoff output is off (wg_set_pot_cell_runoff)");
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  std::vector<double> buf(h->n);
  HIP_CHECK(h, hipMemcpy(buf.data(), h->X.pot_acc + size_t(member) * h->n, h->n * sizeof(double),
                         hipMemcpyDeviceToHost));
  for (int32_t i = 0; i < h->n; i++) pot[i] = buf[h->inv[i]];
  return WG_OK;
}

extern "C" int wg_reset_station_discharge(wg_handle *h) {
  if (!h) return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  if (h->n_stations > 0) {
    HIP_CHECK(h, hipMemsetAsync(h->X.st_sum, 0, size_t(h->m) * h->n_stations * sizeof(double), h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
  }
  return WG_OK;
}

extern "C" int wg_get_day_outputs(wg_handle *h, int32_t member, wg_day_outputs *o) {
  if (!h || !o || member < 0 || member >= h->m) return WG_E_ARG;
  HIP_CHECK(h, hipSetDevice(h->device));
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  const int32_t n = h->n;
  const size_t off = size_t(member) * n;
  std::vector<double> bd(n);
  std::pair<double *, double *> pairs[] = {{o->discharge, h->X.tv},          {o->surface_runoff, h->X.d_surf},
                                           {o->gw_recharge, h->X.d_recharge}, {o->open_water_pet, h->X.d_ow_pet},
                                           {o->cell_aet_uncorr, h->X.d_aet},  {o->prec, h->X.d_prec}};
  for (auto &pr : pairs) {
    if (!pr.first) continue;
    HIP_CHECK(h, hipMemcpy(bd.data(), pr.second + off, n * sizeof(double), hipMemcpyDeviceToHost));
    for (int32_t p = 0; p < n; p++) pr.first[h->perm[p]] = bd[p];
  }
  return WG_OK;
}

extern "C" int wg_synchronize(wg_handle *h) {
  if (!h) return WG_E_ARG;
  HIP_CHECK(h, hipStreamSynchronize(h->stream));
  return WG_OK;
}

extern "C" void *wg_stream(wg_handle *h) { return h ? (void *)h->stream : nullptr; }

extern "C" int wg_last_timing(wg_handle *h, double *ms, int64_t *n_launches) {
  if (!h) return WG_E_ARG;
  for (int i = 0; i < 3; i++) {
    if (ms) ms[i] = h->last_ms[i];
    if (n_launches) n_launches[i] = h->last_launches[i];
  }
  return WG_OK;
}

// Schedule facts for tests and the benchmark: [0] groups, [1] max group depth,
// [2] level slots, [3] cells in level 0 of their group.
extern "C" int wg_schedule_info(wg_handle *h, int32_t *info) {
  if (!h || !info) return WG_E_ARG;
  info[0] = h->n_groups;
  info[1] = h->max_depth;
  info[2] = int32_t(h->level_off.size()) - 1;
  int32_t lvl0 = 0;
  for (int32_t g = 0; g < h->n_groups; g++) {
    const int32_t s = h->group_first_level[g];
    lvl0 += h->level_off[s + 1] - h->level_off[s];
  }
  info[3] = lvl0;
  return WG_OK;
}

extern "C" int wg_partitions(wg_handle *h, int32_t max_parts, int32_t *cells, int32_t *depth) {
  if (!h || max_parts < 0 || (max_parts > 0 && !cells)) return WG_E_ARG;
  const int32_t np = int32_t(h->parts.size());
  for (int32_t pp = 0; pp < np && pp < max_parts; pp++) {
    const WgPart &P = h->parts[pp];
    cells[pp] = P.cell_end - P.cell_begin;
    if (depth) {
      // the deepest group whose cells lie in the partition
      int32_t d = 0;
      for (int32_t g = 0; g < h->n_groups; g++) {
        const int32_t a = h->level_off[h->group_first_level[g]];
        if (a >= P.cell_begin && a < P.cell_end)
          d = std::max(d, h->group_first_level[g + 1] - h->group_first_level[g]);
      }
      depth[pp] = d;
    }
  }
  return np;
}

extern "C" int wg_libm_eval(int32_t fn, int32_t n, const double *x, const double *y, double *out) {
  if (n <= 0 || !x || !out || fn < 0 || fn > 5 || (fn == 1 && !y)) return WG_E_ARG;
  double *dx = nullptr, *dy = nullptr, *dout = nullptr;
  const size_t b = size_t(n) * sizeof(double);
  int rc = WG_OK;
  if (hipMalloc(&dx, b) != hipSuccess || hipMalloc(&dy, b) != hipSuccess || hipMalloc(&dout, b) != hipSuccess) {
    rc = WG_E_NOMEM;
  } else if (hipMemcpy(dx, x, b, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(dy, y ? y : x, b, hipMemcpyHostToDevice) != hipSuccess) {
    rc = WG_E_HIP;
  } else {
    hipLaunchKernelGGL(wg_libm_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, fn, n, dx, dy, dout);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, dout, b, hipMemcpyDeviceToHost) != hipSuccess)
      rc = WG_E_HIP;
  }
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (dout) (void)hipFree(dout);
  return rc;
}
