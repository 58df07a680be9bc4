// wg_world.cpp -- seeded synthetic WaterGAP grids (see wg_world.h).
//
// Geometry follows the reference's 0.5-degree conventions: row 1 is 89.75 N,
// column 1 is 179.75 W (rout_prepare.cpp:217-218), cell areas from the spherical
// zone of each row (GAREA.UNF0), D8 keypad directions (rout_prepare.cpp:79-113),
// river slope and length from the cell distances (rout_prepare.cpp:687-832).
#include "wg_world.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>
#include <queue>
#include <thread>
#include <vector>

namespace {

// Philox4x32-10 (Salmon et al., SC'11).
struct Philox {
  static void block(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
      uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
      uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
      uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0;
      uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1;
      c[0] = n0;
      c[1] = uint32_t(p1);
      c[2] = n2;
      c[3] = uint32_t(p0);
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
  }
};

// Counter-based stream: (seed, stream) key, (a, b) position, draw index.
struct Rng {
  uint32_t k0, k1, a, b, blk = 0;
  uint32_t buf[4];
  int used = 4;
  Rng(uint64_t seed, uint32_t stream, uint32_t a_, uint32_t b_)
      : k0(uint32_t(seed) ^ (stream * 0x9E3779B9u)), k1(uint32_t(seed >> 32) + stream), a(a_), b(b_) {}
  uint32_t u32() {
    if (used == 4) {
      buf[0] = a;
      buf[1] = b;
      buf[2] = blk++;
      buf[3] = 0x57474150u;
      Philox::block(buf, k0, k1);
      used = 0;
    }
    return buf[used++];
  }
  double uniform() { return (u32() + 0.5) * 2.3283064365386963e-10; }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // Marsaglia-Tsang with the u^(1/k) boost for k < 1.
  double gamma(double k, double scale) {
    double boost = 1.0;
    if (k < 1.0) {
      boost = std::pow(uniform(), 1.0 / k);
      k += 1.0;
    }
    const double d = k - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double x = normal(), v = 1.0 + c * x;
      if (v <= 0) continue;
      v = v * v * v;
      double u = uniform();
      if (std::log(u) < 0.5 * x * x + d - d * v + d * std::log(v)) return d * v * boost * scale;
    }
  }
};

constexpr int kRows = 360, kCols = 720;
constexpr double kPi = 3.141592653589793;

double lat_of_row(int row) { return 89.75 - (row - 1) * 0.5; }

// Spherical-zone area of a 0.5 x 0.5 degree cell in row `row`, km2, stored as
// float like GAREA.UNF0.
double row_area(int row) {
  const double R = 6371.0;
  const double top = (90.0 - (row - 1) * 0.5) * kPi / 180.0;
  const double bot = (90.0 - row * 0.5) * kPi / 180.0;
  const double a = R * R * (0.5 * kPi / 180.0) * (std::sin(top) - std::sin(bot));
  return double(float(a));
}

// Value noise with smooth interpolation, periodic in longitude.
struct ValueNoise {
  uint64_t seed;
  uint32_t stream;
  double lattice(int i, int j, int period) const {
    i = ((i % period) + period) % period;
    Rng r(seed, stream, uint32_t(i), uint32_t(j));
    return r.uniform();
  }
  double at(double x, double y, double cell, int period_cells) const {
    const double fx = x / cell, fy = y / cell;
    const int ix = int(std::floor(fx)), iy = int(std::floor(fy));
    const double tx = fx - ix, ty = fy - iy;
    const double sx = tx * tx * (3 - 2 * tx), sy = ty * ty * (3 - 2 * ty);
    const double v00 = lattice(ix, iy, period_cells), v10 = lattice(ix + 1, iy, period_cells);
    const double v01 = lattice(ix, iy + 1, period_cells), v11 = lattice(ix + 1, iy + 1, period_cells);
    return (v00 * (1 - sx) + v10 * sx) * (1 - sy) + (v01 * (1 - sx) + v11 * sx) * sy;
  }
  // fractal sum over octaves; x in [0, width) wraps when wrap_width > 0
  double fbm(double x, double y, double base_cell, int octaves, double wrap_width) const {
    double s = 0, amp = 1, norm = 0, cell = base_cell;
    for (int o = 0; o < octaves; o++) {
      int period = wrap_width > 0 ? std::max(1, int(std::lround(wrap_width / cell))) : 1 << 30;
      double c = wrap_width > 0 ? wrap_width / period : cell;
      ValueNoise oct{seed, stream + 101u * o};
      s += amp * oct.at(x, y, c, period);
      norm += amp;
      amp *= 0.55;
      cell *= 0.5;
    }
    return s / norm;
  }
};

template <class T>
T *alloc(size_t n) {
  return static_cast<T *>(std::calloc(n ? n : 1, sizeof(T)));
}

// Land cover tables in the layout of LCT_22.DAT / LAI_22.DAT (18 classes).
void fill_tables(wg_lct_tables &t) {
  const float root[18] = {4.0f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 1.5f, 1.0f, 1.0f,
                          1.0f, 1.5f, 0.1f, 0.5f, 0.5f, 1.0f, 1.0f, 0.1f, 1.0f};
  const double alb[18] = {0.15, 0.14, 0.18, 0.16, 0.17, 0.19, 0.20, 0.20, 0.23,
                          0.20, 0.18, 0.35, 0.20, 0.22, 0.23, 0.23, 0.14, 0.18};
  const double albs[18] = {0.30, 0.30, 0.45, 0.45, 0.45, 0.55, 0.60, 0.60, 0.70,
                           0.65, 0.60, 0.80, 0.60, 0.65, 0.65, 0.65, 0.50, 0.60};
  const double ddf[18] = {1.5, 1.5, 3.0, 3.0, 3.0, 4.0, 4.0, 4.5, 5.0,
                          5.0, 4.0, 6.0, 4.0, 5.0, 5.0, 5.0, 3.0, 4.0};
  const double emis[18] = {0.99, 0.99, 0.98, 0.98, 0.98, 0.97, 0.96, 0.96, 0.96,
                           0.96, 0.97, 0.92, 0.95, 0.96, 0.96, 0.96, 0.99, 0.97};
  const float lai[18] = {4.8f, 5.0f, 3.6f, 4.0f, 4.6f, 2.5f, 1.7f, 1.6f, 1.7f,
                         1.8f, 2.5f, 0.5f, 1.5f, 2.0f, 3.0f, 3.0f, 1.0f, 2.5f};
  const float decid[18] = {0.0f, 0.0f, 1.0f, 1.0f, 0.5f, 0.8f, 0.5f, 1.0f, 1.0f,
                           1.0f, 0.6f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 0.8f};
  const float everg[18] = {1.0f, 1.0f, 0.0f, 0.0f, 0.5f, 0.2f, 0.5f, 0.0f, 0.0f,
                           0.0f, 0.4f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.2f};
  const int16_t idays[18] = {10, 10, 10, 10, 10, 20, 20, 10, 10, 10, 10, 10, 20, 10, 30, 30, 10, 10};
  for (int i = 0; i < 18; i++) {
    t.rooting_depth[i] = root[i];
    t.albedo[i] = alb[i];
    t.albedo_snow[i] = albs[i];
    t.ddf[i] = ddf[i];
    t.emissivity[i] = emis[i];
    t.lai[i] = lai[i];
    t.decid_frac[i] = decid[i];
    t.evergreen_red[i] = everg[i];
    t.initial_days[i] = idays[i];
    t.kc_min[i] = 0.3 + 0.01 * i;
    t.kc_max[i] = 1.0 + 0.01 * i;
  }
}

// Default values of the 26 calibration parameters (calib_param.h:72-98).
double param_default(int p) {
  switch (p) {
    case WG_P_GAMRUN_C: return 2.0;
    case WG_P_CFA: return 1.0;
    case WG_P_CFS: return 1.0;
    case WG_P_LAK_D: return 5.0;
    case WG_P_WET_D: return 2.0;
    case WG_P_SWOUTF_C: return 0.01;
    case WG_P_PTC_HUM: return 1.26;
    case WG_P_PTC_ARI: return 1.74;
    case WG_P_PET_MXDY: return 15.0;
    case WG_P_MCWH: return 0.3;
    case WG_P_T_SNOWFZ: return 0.0;
    case WG_P_T_SNOWMT: return 0.0;
    case WG_P_T_GRADNT: return 0.006;
    case WG_P_PCRITGWA: return 12.5;
    case WG_P_GWOUTF_C: return 0.01;
    default: return 1.0;  // multipliers
  }
}

// Cell-distance table of rout_prepare.cpp:690-742 (float).
struct CellDist {
  float d[9][360];
  CellDist() {
    const float pi = 3.141592653589793f, earth_radius = 6371.211f;
    float vert = 2 * pi * earth_radius * 0.5f / 360.0f;
    for (int i = 0; i < 360; i++) {
      float l = 0.25f + i * 0.5f;
      float horiz = 2 * pi * earth_radius * std::sin(l * pi / 180.0f) * 0.5f / 360.0f;
      d[1][i] = vert;
      d[3][i] = horiz;
      d[4][i] = 0;
      d[5][i] = horiz;
      d[7][i] = vert;
    }
    for (int i = 0; i < 359; i++) d[0][i] = std::sqrt(d[3][i] * d[3][i + 1] + d[1][i] * d[1][i]);
    d[0][359] = -99;
    for (int i = 1; i < 360; i++) d[6][i] = std::sqrt(d[3][i] * d[3][i - 1] + d[1][i] * d[1][i]);
    d[6][0] = -99;
    for (int i = 0; i < 360; i++) {
      d[2][i] = d[0][i];
      d[8][i] = d[6][i];
    }
  }
};

const int kDc[10] = {0, -1, 0, 1, -1, 0, 1, -1, 0, 1};
const int kDr[10] = {0, 1, 1, 1, 0, 0, 0, -1, -1, -1};

struct Builder {
  const wg_world_spec &sp;
  wg_world *w;
  std::vector<int> cells_r, cells_c;   // 1-based rows/cols of land cells
  std::vector<float> dem;              // raw elevation per cell

  explicit Builder(const wg_world_spec &s) : sp(s), w(alloc<wg_world>(1)) {}

  int32_t at(int col, int row) const {
    if (row < 1 || row > kRows) return 0;
    if (col < 1) col += kCols;
    if (col > kCols) col -= kCols;
    return w->cell_at[(row - 1) * kCols + (col - 1)];
  }

  void make_mask() {
    w->cell_at = alloc<int32_t>(size_t(kRows) * kCols);
    if (sp.kind == 0) {
      for (int r = 0; r < sp.nrows; r++)
        for (int c = 0; c < sp.ncols; c++) {
          cells_r.push_back(sp.row0 + r);
          cells_c.push_back(sp.col0 + c);
        }
      // diamond-square style fractal relief over the lattice
      ValueNoise vn{sp.seed, 11};
      for (size_t i = 0; i < cells_r.size(); i++) {
        double v = vn.fbm(cells_c[i] - sp.col0 + 0.37, cells_r[i] - sp.row0 + 0.61, 16.0, 5, 0);
        dem.push_back(float(v));
      }
    } else {
      // global mask: the n_land highest points of a fractal field, poles excluded
      ValueNoise vn{sp.seed, 21};
      std::vector<std::pair<double, int>> score;
      score.reserve(size_t(kRows) * kCols);
      for (int r = 1; r <= kRows; r++) {
        double lat = lat_of_row(r);
        if (std::fabs(lat) > 80.0) continue;
        double pen = 0.25 * std::pow(std::fabs(lat) / 80.0, 4.0);
        for (int c = 1; c <= kCols; c++) {
          double v = vn.fbm(c - 0.5, r - 0.5, 90.0, 6, kCols) - pen;
          score.push_back({-v, (r - 1) * kCols + (c - 1)});
        }
      }
      std::sort(score.begin(), score.end());
      std::vector<int> pos;
      for (int i = 0; i < sp.n_land && i < int(score.size()); i++) pos.push_back(score[i].second);
      std::sort(pos.begin(), pos.end());  // cells numbered row-major from the north (GCRC order)
      ValueNoise rel{sp.seed, 31};
      for (int p : pos) {
        int r = p / kCols + 1, c = p % kCols + 1;
        cells_r.push_back(r);
        cells_c.push_back(c);
        double v = rel.fbm(c - 0.5, r - 0.5, sp.relief_scale > 0 ? sp.relief_scale : 40.0, 6, kCols);
        dem.push_back(float(v));
      }
    }
    w->n = int32_t(cells_r.size());
    for (int i = 0; i < w->n; i++) w->cell_at[(cells_r[i] - 1) * kCols + (cells_c[i] - 1)] = i + 1;
  }

  // Priority flood from the outlet seeds: every cell drains to the neighbour it
  // was reached from; the filled DEM is the pit-filled altitude.
  void flow_directions() {
    const int n = w->n;
    w->ldd = alloc<int8_t>(n);
    w->altitude = alloc<float>(n);
    std::vector<double> elev(n);
    // rescale relief to 0..3000 m, plus a rise away from the coast for the global mask
    double lo = 1e30, hi = -1e30;
    for (int i = 0; i < n; i++) {
      lo = std::min(lo, double(dem[i]));
      hi = std::max(hi, double(dem[i]));
    }
    for (int i = 0; i < n; i++) elev[i] = 3000.0 * (dem[i] - lo) / std::max(hi - lo, 1e-9);

    std::vector<int> seeds;
    std::vector<int8_t> seed_ldd;
    std::vector<char> coastal(n, 0);
    for (int i = 0; i < n; i++)
      for (int k = 1; k <= 9; k++)
        if (k != 5 && !at(cells_c[i] + kDc[k], cells_r[i] + kDr[k])) coastal[i] = 1;
    Rng rng(sp.seed, 41, 0, 0);
    auto ocean_dir = [&](int i) -> int8_t {
      for (int k = 1; k <= 9; k++)
        if (k != 5 && !at(cells_c[i] + kDc[k], cells_r[i] + kDr[k])) return int8_t(k);
      return 5;
    };
    if (sp.kind == 0) {
      int best = -1;
      for (int i = 0; i < n; i++)
        if (coastal[i] && (best < 0 || elev[i] < elev[best])) best = i;
      seeds.push_back(best);
      seed_ldd.push_back(ocean_dir(best));
    } else {
      // connected components (8-neighbourhood); each gets its lowest coastal cell as outlet
      std::vector<int> comp(n, -1);
      int ncomp = 0;
      for (int s = 0; s < n; s++) {
        if (comp[s] >= 0) continue;
        std::vector<int> st{s};
        comp[s] = ncomp;
        int best = -1;
        while (!st.empty()) {
          int c = st.back();
          st.pop_back();
          if (coastal[c] && (best < 0 || elev[c] < elev[best])) best = c;
          for (int k = 1; k <= 9; k++) {
            if (k == 5) continue;
            int g = at(cells_c[c] + kDc[k], cells_r[c] + kDr[k]);
            if (g && comp[g - 1] < 0) {
              comp[g - 1] = ncomp;
              st.push_back(g - 1);
            }
          }
        }
        if (best < 0) best = s;
        seeds.push_back(best);
        seed_ldd.push_back(coastal[best] ? ocean_dir(best) : int8_t(5));
        ncomp++;
      }
      std::vector<char> is_seed(n, 0);
      for (int s : seeds) is_seed[s] = 1;
      std::vector<int> coast;
      for (int i = 0; i < n; i++)
        if (coastal[i] && !is_seed[i]) coast.push_back(i);
      // extra coastal outlets, lowest-first among a random half
      for (int i = int(coast.size()) - 1; i > 0; i--) std::swap(coast[i], coast[rng.u32() % (i + 1)]);
      int extra = std::max(0, sp.n_outlets - int(seeds.size()));
      for (int j = 0; j < extra && j < int(coast.size()); j++) {
        seeds.push_back(coast[j]);
        seed_ldd.push_back(ocean_dir(coast[j]));
        is_seed[coast[j]] = 1;
      }
      // inland sinks: interior cells in the upper elevation range become closed basins
      std::vector<int> interior;
      for (int i = 0; i < n; i++)
        if (!coastal[i]) interior.push_back(i);
      for (int j = 0; j < sp.n_sinks && !interior.empty(); j++) {
        int pick = interior[rng.u32() % interior.size()];
        if (is_seed[pick]) continue;
        is_seed[pick] = 1;
        seeds.push_back(pick);
        seed_ldd.push_back(-1);
        elev[pick] *= 0.5;
      }
    }
    // flood
    std::vector<double> filled(n, 0);
    std::vector<char> done(n, 0);
    using Item = std::pair<std::pair<double, int64_t>, int>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    int64_t counter = 0;
    for (size_t s = 0; s < seeds.size(); s++) {
      int c = seeds[s];
      done[c] = 1;
      filled[c] = elev[c];
      w->ldd[c] = seed_ldd[s];
      pq.push({{filled[c], counter++}, c});
    }
    while (!pq.empty()) {
      int c = pq.top().second;
      pq.pop();
      for (int k = 1; k <= 9; k++) {
        if (k == 5) continue;
        int g = at(cells_c[c] + kDc[k], cells_r[c] + kDr[k]);
        if (!g || done[g - 1]) continue;
        int j = g - 1;
        done[j] = 1;
        filled[j] = std::max(elev[j], filled[c]);
        w->ldd[j] = int8_t(10 - k);  // neighbour at keypad k drains back in direction 10-k
        pq.push({{filled[j], counter++}, j});
      }
    }
    for (int i = 0; i < n; i++) w->altitude[i] = float(filled[i]);
  }

  void schedule() {
    const int n = w->n;
    w->row = alloc<int16_t>(n);
    w->col = alloc<int16_t>(n);
    for (int i = 0; i < n; i++) {
      w->row[i] = int16_t(cells_r[i]);
      w->col[i] = int16_t(cells_c[i]);
    }
    w->inflow9 = alloc<int32_t>(size_t(n) * 9);
    w->downstream = alloc<int32_t>(n);
    w->rout_order = alloc<int32_t>(n);
    w->level = alloc<int32_t>(n);
    wg_inflow_outflow(n, w->ldd, w->row, w->col, w->cell_at, w->inflow9, w->downstream);
    wg_rout_order(n, w->inflow9, w->downstream, w->rout_order, w->level);
    w->max_level = 0;
    for (int i = 0; i < n; i++) w->max_level = std::max(w->max_level, w->level[i]);
    // basins: follow downstream links to the outlet
    std::vector<int> by_order(n);
    for (int i = 0; i < n; i++) by_order[w->rout_order[i] - 1] = i;
    w->basin = alloc<int32_t>(n);
    std::vector<int> outlet_of(n, -1);
    for (int k = n - 1; k >= 0; k--) {
      int i = by_order[k];
      int d = w->downstream[i];
      outlet_of[i] = d > 0 ? outlet_of[d - 1] : i;
    }
    std::vector<int> id(n, -1);
    std::vector<int> outlets;
    for (int i = 0; i < n; i++) {
      int o = outlet_of[i];
      if (id[o] < 0) {
        id[o] = int(outlets.size());
        outlets.push_back(o);
      }
      w->basin[i] = id[o];
    }
    w->n_basins = int32_t(outlets.size());
    w->basin_outlet = alloc<int32_t>(outlets.size());
    for (size_t b = 0; b < outlets.size(); b++) w->basin_outlet[b] = outlets[b];
    // upstream area in routing order
    w->cell_area = alloc<double>(n);
    w->upstream_area = alloc<double>(n);
    for (int i = 0; i < n; i++) w->cell_area[i] = row_area(w->row[i]);
    for (int k = 0; k < n; k++) {
      int i = by_order[k];
      w->upstream_area[i] += w->cell_area[i];
      if (w->downstream[i] > 0) w->upstream_area[w->downstream[i] - 1] += w->upstream_area[i];
    }
  }

  void statics() {
    const int n = w->n;
    fill_tables(w->lct);
    w->contfreq = alloc<double>(n);
    w->aindex = alloc<int16_t>(n);
    w->land_cover = alloc<int8_t>(n);
    w->texture = alloc<int8_t>(n);
    w->built_up = alloc<float>(n);
    w->smax = alloc<float>(n);
    w->gw_factor = alloc<float>(n);
    w->rgmax = alloc<int16_t>(n);
    w->elevation = alloc<int16_t>(size_t(n) * WG_NELEV);
    w->params = alloc<double>(size_t(WG_NPARAM) * n);
    w->glo_lake = alloc<double>(n);
    w->loc_lake = alloc<double>(n);
    w->glo_wetland = alloc<double>(n);
    w->loc_wetland = alloc<double>(n);
    w->lake_area = alloc<double>(n);
    w->reservoir_area = alloc<double>(n);
    w->glo_res = alloc<double>(n);
    w->stor_cap = alloc<double>(n);
    w->mean_outflow = alloc<double>(n);
    w->res_type = alloc<int8_t>(n);
    w->start_month = alloc<int8_t>(n);
    w->river_length = alloc<float>(n);
    w->river_slope = alloc<float>(n);
    w->roughness = alloc<float>(n);
    w->bankfull_flow = alloc<float>(n);
    w->meander = alloc<float>(n);
    static const CellDist cd;
    for (int i = 0; i < n; i++) {
      Rng r(sp.seed, 51, uint32_t(i), 0);
      const double lat = lat_of_row(w->row[i]);
      bool coastal = false;
      for (int k = 1; k <= 9; k++)
        if (k != 5 && !at(w->col[i] + kDc[k], w->row[i] + kDr[k])) coastal = true;
      w->contfreq[i] = coastal ? double(float(40.0 + 60.0 * r.uniform())) : 100.0;
      const bool dry = (sp.kind == 1) ? (std::fabs(std::fabs(lat) - 25.0) < 12.0 && r.uniform() < 0.7)
                                      : (r.uniform() < 0.3);
      w->aindex[i] = dry ? 1 : 0;
      w->land_cover[i] = int8_t(1 + r.u32() % 18);
      w->texture[i] = int8_t(10 + r.u32() % 20);
      w->built_up[i] = (r.uniform() < 0.05) ? float(0.3 * r.uniform()) : 0.0f;
      w->smax[i] = float(50.0 + 450.0 * r.uniform());
      w->gw_factor[i] = float(0.1 + 0.8 * r.uniform());
      double rg = 1.0 + 6.0 * r.uniform();
      w->rgmax[i] = int16_t(std::floor(rg * 100 + 0.5));
      for (int p = 0; p < WG_NPARAM; p++) w->params[size_t(p) * n + i] = param_default(p);
      w->params[size_t(WG_P_GAMRUN_C) * n + i] = 0.1 + 4.9 * r.uniform();
      // elevation sub-cells: cell mean +- lognormal(0.5) spread, >= 1 m, ascending
      const double zbar = std::max(1.0, double(w->altitude[i]));
      int16_t *ev = w->elevation + size_t(i) * WG_NELEV;
      ev[0] = int16_t(std::lround(zbar));
      std::vector<int> z(100);
      const double h = 20.0 + 0.1 * zbar;
      for (int j = 0; j < 100; j++) {
        double d = h * std::exp(0.5 * r.normal());
        if (r.uniform() < 0.5) d = -d;
        z[j] = int(std::lround(std::min(std::max(zbar + d, 1.0), 8800.0)));
      }
      std::sort(z.begin(), z.end());
      for (int j = 0; j < 100; j++) ev[j + 1] = int16_t(z[j]);
      // surface water bodies (%)
      const double wf = 100.0 * sp.water_frac_max;
      w->loc_lake[i] = double(float(wf * r.uniform()));
      w->loc_wetland[i] = double(float(wf * r.uniform()));
      w->glo_wetland[i] = double(float(wf * r.uniform()));
      if (sp.glolake_every > 0 && r.u32() % uint32_t(sp.glolake_every) == 0) {
        w->glo_lake[i] = double(float(1.0 + 19.0 * r.uniform()));
        w->lake_area[i] = double(float(w->glo_lake[i] / 100.0 * w->cell_area[i]));
      }
      if (sp.res_every > 0 && r.u32() % uint32_t(sp.res_every) == 0) {
        w->glo_res[i] = double(float(1.0 + 9.0 * r.uniform()));
        w->reservoir_area[i] = double(float(w->glo_res[i] / 100.0 * w->cell_area[i]));
        w->stor_cap[i] = double(float(w->reservoir_area[i] * (5.0 + 25.0 * r.uniform()) / 1000.0));
        // mean annual inflow from 250 mm/yr runoff over the upstream area, m3/s
        w->mean_outflow[i] = double(float(w->upstream_area[i] * 1.0e6 * 0.25 / 31536000.0));
        w->res_type[i] = int8_t(1 + (r.u32() & 1));
        w->start_month[i] = int8_t(1 + r.u32() % 12);
      }
      // river geometry (rout_prepare.cpp:761-832), roughness and bankfull flow
      w->meander[i] = float(1.0 + 0.5 * r.uniform());
      w->roughness[i] = float(0.03 + 0.05 * r.uniform());
      double qmean = w->upstream_area[i] * 1.0e6 * 0.25 / 31536000.0;
      w->bankfull_flow[i] = float(std::max(0.05, 2.0 * qmean));
    }
    // slope and length need the meandering ratio of the outflow cell
    for (int i = 0; i < n; i++) {
      const int o = w->downstream[i] - 1;
      const int l = w->ldd[i];
      if (o >= 0 && l != 5 && l != -1) {
        float s = (w->altitude[i] - w->altitude[o]) /
                  (1000.0f * cd.d[l - 1][w->row[i] - 1] * ((w->meander[i] + w->meander[o]) / 2));
        w->river_slope[i] = s;
      } else {
        w->river_slope[i] = 0;
      }
      if (w->river_slope[i] < 0.00001f) w->river_slope[i] = 0.00001f;
      float len;
      if (o == -1)
        len = 55.f;
      else if (w->row[i] == w->row[o])
        len = cd.d[3][w->row[i] - 1];
      else if (w->row[i] > w->row[o])
        len = (w->col[i] == w->col[o]) ? cd.d[7][w->row[i] - 1] : cd.d[8][w->row[i] - 1];
      else
        len = (w->col[i] == w->col[o]) ? cd.d[1][w->row[i] - 1] : cd.d[2][w->row[i] - 1];
      if (o == -1)
        len *= w->meander[i];
      else
        len *= ((w->meander[i] + w->meander[o]) / 2);
      w->river_length[i] = len;
    }
    if (sp.edge_cases) edge_cases();
  }

  // Cells that take rarely used branches: land area fraction driven to zero,
  // missing soil capacity, arid cells with coarse texture, built-up cells, CFA != 1.
  void edge_cases() {
    const int n = w->n;
    for (int i = 0; i < n; i++) {
      Rng r(sp.seed, 61, uint32_t(i), 0);
      const uint32_t pick = r.u32() % 50;
      if (pick == 0) {
        // continental area exactly covered by water bodies -> land area fraction 0
        w->contfreq[i] = w->glo_lake[i] + w->glo_wetland[i] + w->loc_lake[i] + w->loc_wetland[i] + w->glo_res[i];
        if (w->contfreq[i] <= 0) w->contfreq[i] = 1.0, w->loc_wetland[i] = 1.0;
      } else if (pick == 1) {
        w->smax[i] = -9999.0f;
      } else if (pick == 2) {
        w->aindex[i] = 1;
        w->texture[i] = 25;
      } else if (pick == 3) {
        w->built_up[i] = 0.5f;
      } else if (pick == 4 || pick == 5) {
        w->params[size_t(WG_P_CFA) * n + i] = 0.5 + r.uniform();
      } else if (pick == 6) {
        w->loc_lake[i] = w->loc_wetland[i] = w->glo_wetland[i] = 0.0;
      }
    }
  }
};

}  // namespace

extern "C" int wg_world_create(const wg_world_spec *spec, wg_world **out) {
  if (!spec || !out) return WG_E_ARG;
  if (spec->kind == 0 && (spec->nrows <= 0 || spec->ncols <= 0 || spec->row0 < 1 ||
                          spec->row0 + spec->nrows - 1 > kRows || spec->col0 < 1 ||
                          spec->col0 + spec->ncols - 1 > kCols))
    return WG_E_ARG;
  if (spec->kind == 1 && (spec->n_land <= 0 || spec->n_land > 200000)) return WG_E_ARG;
  Builder b(*spec);
  b.w->kind = spec->kind;
  b.w->seed = spec->seed;
  b.make_mask();
  b.flow_directions();
  b.schedule();
  b.statics();
  *out = b.w;
  return WG_OK;
}

extern "C" void wg_world_destroy(wg_world *w) {
  if (!w) return;
  void *p[] = {w->row, w->col, w->cell_at, w->cell_area, w->contfreq, w->aindex, w->ldd,
               w->land_cover, w->texture, w->res_type, w->start_month, w->built_up, w->smax,
               w->gw_factor, w->rgmax, w->elevation, w->params, w->glo_lake, w->loc_lake,
               w->glo_wetland, w->loc_wetland, w->lake_area, w->reservoir_area, w->glo_res,
               w->stor_cap, w->mean_outflow, w->river_length, w->river_slope, w->roughness,
               w->bankfull_flow, w->altitude, w->meander, w->inflow9, w->downstream,
               w->rout_order, w->level, w->basin, w->basin_outlet, w->upstream_area};
  for (void *q : p) std::free(q);
  std::free(w);
}

extern "C" void wg_world_statics(const wg_world *w, wg_static_fields *s) {
  std::memset(s, 0, sizeof(*s));
  s->row = w->row;
  s->cell_area = w->cell_area;
  s->contfreq = w->contfreq;
  s->aindex = w->aindex;
  s->ldd = w->ldd;
  s->land_cover = w->land_cover;
  s->built_up = w->built_up;
  s->smax = w->smax;
  s->gw_factor = w->gw_factor;
  s->rgmax = w->rgmax;
  s->texture = w->texture;
  s->elevation = w->elevation;
  s->params = w->params;
  s->glo_lake = w->glo_lake;
  s->loc_lake = w->loc_lake;
  s->glo_wetland = w->glo_wetland;
  s->loc_wetland = w->loc_wetland;
  s->lake_area = w->lake_area;
  s->reservoir_area = w->reservoir_area;
  s->glo_res = w->glo_res;
  s->stor_cap = w->stor_cap;
  s->mean_outflow = w->mean_outflow;
  s->res_type = w->res_type;
  s->start_month = w->start_month;
  s->river_length = w->river_length;
  s->river_slope = w->river_slope;
  s->roughness = w->roughness;
  s->bankfull_flow = w->bankfull_flow;
  s->downstream = w->downstream;
  s->rout_order = w->rout_order;
  s->lct = &w->lct;
}

extern "C" int wg_world_forcing(const wg_world *w, uint64_t seed, int64_t day0, int32_t n_days, float *out) {
  if (!w || !out || n_days < 0) return WG_E_ARG;
  const int n = w->n;
  // cells are independent streams: split them over threads
  auto work = [&](int c0, int c1) {
    for (int32_t d = 0; d < n_days; d++) {
      const int64_t gd = day0 + d;
      const int doy = int(gd % 365) + 1;
      float *T = out + size_t(d) * 4 * n, *P = T + n, *SW = P + n, *LW = SW + n;
      const double season = std::sin(2.0 * kPi * (doy - 100) / 365.0);
      for (int i = c0; i < c1; i++) {
        Rng r(seed, 71, uint32_t(i), uint32_t(gd));
        const double lat = lat_of_row(w->row[i]);
        double tmean = 10.0, amp = 12.0, s = season;
        if (w->kind == 1) {
          tmean = 28.0 - 0.55 * std::fabs(lat);
          amp = 2.0 + 0.25 * std::fabs(lat);
          if (lat < 0) s = -season;
        }
        T[i] = float(tmean + amp * s + 3.0 * r.normal());
        P[i] = (r.uniform() < 0.4) ? float(r.gamma(0.8, 8.0)) : 0.0f;
        SW[i] = float(std::min(300.0, std::max(50.0, 175.0 + 125.0 * s + 15.0 * r.normal())));
        LW[i] = float(std::min(400.0, std::max(250.0, 325.0 + 75.0 * s + 10.0 * r.normal())));
      }
    }
  };
  int nt = int(std::thread::hardware_concurrency());
  if (const char *v = std::getenv("WG_WORLD_THREADS")) nt = std::atoi(v);
  nt = std::max(1, std::min(nt, 16));
  if (int64_t(n) * n_days < 200000) nt = 1;
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++) {
    const int c0 = int(int64_t(n) * t / nt), c1 = int(int64_t(n) * (t + 1) / nt);
    pool.emplace_back(work, c0, c1);
  }
  for (auto &th : pool) th.join();
  return WG_OK;
}

extern "C" int wg_world_subset(const wg_world *w, int32_t m, const int32_t *cells, wg_world **out) {
  if (!w || !cells || !out || m <= 0 || m > w->n) return WG_E_ARG;
  const int n = w->n;
  std::vector<int32_t> newid(n, 0);
  for (int32_t k = 0; k < m; k++) {
    if (cells[k] < 0 || cells[k] >= n || newid[cells[k]]) return WG_E_ARG;
    newid[cells[k]] = k + 1;
  }
  wg_world *s = alloc<wg_world>(1);
  s->n = m;
  s->kind = w->kind;
  s->seed = w->seed;
  s->lct = w->lct;
#define SUB(field, T)                                                       \
  s->field = alloc<T>(m);                                                   \
  for (int32_t k = 0; k < m; k++) s->field[k] = w->field[cells[k]];
  SUB(row, int16_t) SUB(col, int16_t) SUB(cell_area, double) SUB(contfreq, double)
  SUB(aindex, int16_t) SUB(ldd, int8_t) SUB(land_cover, int8_t) SUB(texture, int8_t)
  SUB(res_type, int8_t) SUB(start_month, int8_t) SUB(built_up, float) SUB(smax, float)
  SUB(gw_factor, float) SUB(rgmax, int16_t) SUB(glo_lake, double) SUB(loc_lake, double)
  SUB(glo_wetland, double) SUB(loc_wetland, double) SUB(lake_area, double)
  SUB(reservoir_area, double) SUB(glo_res, double) SUB(stor_cap, double) SUB(mean_outflow, double)
  SUB(river_length, float) SUB(river_slope, float) SUB(roughness, float) SUB(bankfull_flow, float)
  SUB(altitude, float) SUB(meander, float) SUB(upstream_area, double)
#undef SUB
  s->elevation = alloc<int16_t>(size_t(m) * WG_NELEV);
  s->params = alloc<double>(size_t(WG_NPARAM) * m);
  for (int32_t k = 0; k < m; k++) {
    std::memcpy(s->elevation + size_t(k) * WG_NELEV, w->elevation + size_t(cells[k]) * WG_NELEV,
                WG_NELEV * sizeof(int16_t));
    for (int p = 0; p < WG_NPARAM; p++) s->params[size_t(p) * m + k] = w->params[size_t(p) * n + cells[k]];
  }
  // lattice restricted to the subset; downstream links leaving the subset become outlets
  s->cell_at = alloc<int32_t>(size_t(kRows) * kCols);
  for (int32_t k = 0; k < m; k++) s->cell_at[(s->row[k] - 1) * kCols + (s->col[k] - 1)] = k + 1;
  s->inflow9 = alloc<int32_t>(size_t(m) * 9);
  s->downstream = alloc<int32_t>(m);
  s->rout_order = alloc<int32_t>(m);
  s->level = alloc<int32_t>(m);
  wg_inflow_outflow(m, s->ldd, s->row, s->col, s->cell_at, s->inflow9, s->downstream);
  int rc = wg_rout_order(m, s->inflow9, s->downstream, s->rout_order, s->level);
  s->max_level = 0;
  for (int32_t k = 0; k < m; k++) s->max_level = std::max(s->max_level, s->level[k]);
  std::vector<int> by_order(m);
  for (int32_t k = 0; k < m; k++) by_order[s->rout_order[k] - 1] = k;
  std::vector<int> outlet_of(m, -1), id(m, -1), outlets;
  for (int k = m - 1; k >= 0; k--) {
    int i = by_order[k];
    int d = s->downstream[i];
    outlet_of[i] = d > 0 ? outlet_of[d - 1] : i;
  }
  s->basin = alloc<int32_t>(m);
  for (int32_t k = 0; k < m; k++) {
    int o = outlet_of[k];
    if (id[o] < 0) {
      id[o] = int(outlets.size());
      outlets.push_back(o);
    }
    s->basin[k] = id[o];
  }
  s->n_basins = int32_t(outlets.size());
  s->basin_outlet = alloc<int32_t>(outlets.size());
  for (size_t b = 0; b < outlets.size(); b++) s->basin_outlet[b] = outlets[b];
  *out = s;
  return rc;
}
