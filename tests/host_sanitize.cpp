// host_sanitize.cpp -- exercises the library's host C++ and the oracle under
// AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Linked from the host sources only (world generator, routing schedule,
// preprocessing, input files, restart files, calibration search) and the oracle's
// C, all built with -fsanitize=address,undefined by g++/gcc; no HIP code is in the
// program.  Any sanitizer report aborts with a non-zero status
// (-fno-sanitize-recover).  Built by `make -C watergap2_amd/csrc sanitize`.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <unistd.h>

#include "watergap2_amd.h"
#include "../watergap2_amd/csrc/wg_world.h"
#include "../oracle/wg_oracle.h"

namespace {

int failures = 0;

void expect(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", what);
    ++failures;
  }
}

#define CALL(expr)                                                                  \
  do {                                                                              \
    int rc_ = (expr);                                                               \
    if (rc_ != 0) {                                                                 \
      std::fprintf(stderr, "FAIL: %s returned %d\n", #expr, rc_);                   \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

wg_world_spec regional(uint64_t seed, int edge) {
  wg_world_spec s;
  std::memset(&s, 0, sizeof s);
  s.kind = 0;
  s.nrows = 30;
  s.ncols = 40;
  s.row0 = 81;
  s.col0 = 361;
  s.seed = seed;
  s.water_frac_max = 0.1;
  s.glolake_every = 20;
  s.res_every = 25;
  s.edge_cases = edge;
  return s;
}

wg_world_spec small_global() {
  wg_world_spec s;
  std::memset(&s, 0, sizeof s);
  s.kind = 1;
  s.n_land = 6000;
  s.n_outlets = 60;
  s.n_sinks = 4;
  s.seed = 3;
  s.water_frac_max = 0.1;
  s.glolake_every = 100;
  s.res_every = 200;
  return s;
}

struct StateBuf {
  std::vector<double> d[23];
  std::vector<double> snow_elev;
  std::vector<int32_t> lai_days, lai_status;
  wg_state s;
  explicit StateBuf(int n) {
    for (auto &v : d) v.assign(n, 0.0);
    snow_elev.assign(size_t(n) * 101, 0.0);
    lai_days.assign(n, 0);
    lai_status.assign(n, 0);
    std::memset(&s, 0, sizeof s);
    double **p[] = {&s.canopy, &s.soil, &s.snow, &s.storage_transfer, &s.lai_precsum, &s.groundwater,
                    &s.loc_lake, &s.loc_wetl, &s.glo_lake, &s.glo_wetl, &s.reservoir, &s.river,
                    &s.red_loc_lake, &s.red_loc_wetl, &s.red_glo_lake, &s.red_glo_wetl, &s.red_res,
                    &s.red_river, &s.river_frac_next, &s.fswb_frac_next, &s.laf, &s.laf_prev,
                    &s.k_release};
    for (int k = 0; k < 23; ++k) *p[k] = d[k].data();
    s.snow_elev = snow_elev.data();
    s.lai_days = lai_days.data();
    s.lai_status = lai_status.data();
  }
};

// World generation, schedule, preprocessing, restart files and the oracle's day
// loop on one grid.
int exercise(const wg_world_spec &spec, int days, const char *label) {
  wg_world *w = nullptr;
  CALL(wg_world_create(&spec, &w));
  const int32_t n = w->n;

  std::vector<int32_t> inflow9(size_t(n) * 9), outflow(n), order(n), level(n);
  CALL(wg_inflow_outflow(n, w->ldd, w->row, w->col, w->cell_at, inflow9.data(), outflow.data()));
  CALL(wg_rout_order(n, inflow9.data(), outflow.data(), order.data(), level.data()));

  // preprocessing from the grid's flow directions
  std::vector<int8_t> ldd_in(w->ldd, w->ldd + n);
  CALL(wg_replace_flow_to_ocean(n, w->row, w->col, w->cell_at, ldd_in.data()));
  wg_routing_inputs in;
  std::memset(&in, 0, sizeof in);
  in.n_cells = n;
  in.ldd = ldd_in.data();
  in.row = w->row;
  in.col = w->col;
  in.cell_at = w->cell_at;
  in.altitude = w->altitude;
  in.meander = w->meander;
  std::vector<double> monthly(size_t(n) * 12);
  for (int32_t i = 0; i < n; ++i)
    for (int m = 0; m < 12; ++m) monthly[size_t(i) * 12 + m] = w->mean_outflow[i] * (1.0 + 0.3 * std::sin(0.5 * m + i));
  in.reservoir_area = w->reservoir_area;
  in.mean_outflow = w->mean_outflow;
  in.mean_outflow_monthly = monthly.data();
  std::vector<int8_t> ldd_out(n), start_month(n);
  std::vector<int32_t> inflow_out(size_t(n) * 9), outflow_out(n), order_out(n);
  std::vector<int16_t> acc(n);
  std::vector<uint16_t> basins(n), to_outlet(n);
  std::vector<float> dist(9 * 360), slope(n), length(n);
  std::vector<double> alloc(size_t(n) * 5);
  wg_routing_files out;
  std::memset(&out, 0, sizeof out);
  out.ldd = ldd_out.data();
  out.inflow9 = inflow_out.data();
  out.outflow = outflow_out.data();
  out.flow_acc = acc.data();
  out.basins = basins.data();
  out.cells_to_outlet = to_outlet.data();
  out.cell_dist = dist.data();
  out.river_slope = slope.data();
  out.river_length = length.data();
  out.rout_order = order_out.data();
  out.alloc_coeff = alloc.data();
  out.start_month = start_month.data();
  CALL(wg_prepare_routing(&in, &out));
  expect(out.n_basins > 0, "basins found");

  // the oracle's day loop
  wg_static_fields st;
  std::memset(&st, 0, sizeof st);
  wg_world_statics(w, &st);
  wgo_model *m = wgo_create(n, &st);
  expect(m != nullptr, "oracle model");
  if (!m) return 1;
  wgo_set_stations(m, w->n_basins, w->basin_outlet);
  std::vector<float> forcing(size_t(days) * 4 * n);
  CALL(wg_world_forcing(w, 7, 0, days, forcing.data()));
  wgo_run(m, forcing.data(), days, 1);
  StateBuf sb(n);
  CALL(wgo_get_state(m, &sb.s));
  for (int32_t i = 0; i < n; ++i) expect(std::isfinite(sb.s.soil[i]) && std::isfinite(sb.s.river[i]), "finite state");
  std::vector<double> q(w->n_basins);
  wgo_station_discharge(m, q.data());

  // restart files: state -> file values -> text files -> values -> state
  std::vector<double> stor(size_t(n) * WG_NSTORE), snow(size_t(n) * 101), add(size_t(n) * WG_NADDITIONAL);
  CALL(wg_state_to_files(n, &st, &sb.s, stor.data(), snow.data(), add.data()));
  char tmpl[] = "/tmp/wg_host_sanitize_XXXXXX";
  char *dir = mkdtemp(tmpl);
  expect(dir != nullptr, "temporary directory");
  if (dir) {
    std::string p1 = std::string(dir) + "/wghm_state.txt", p2 = std::string(dir) + "/snow.txt",
                p3 = std::string(dir) + "/additional.txt", p4 = std::string(dir) + "/G_SOIL.UNF0";
    CALL(wg_write_wghm_state_txt(p1.c_str(), n, stor.data()));
    CALL(wg_write_snow_elevation_txt(p2.c_str(), n, snow.data()));
    CALL(wg_write_additional_txt(p3.c_str(), n, add.data()));
    std::vector<double> stor2(stor.size()), snow2(snow.size()), add2(add.size());
    CALL(wg_read_wghm_state_txt(p1.c_str(), n, stor2.data()));
    CALL(wg_read_snow_elevation_txt(p2.c_str(), n, snow2.data()));
    CALL(wg_read_additional_txt(p3.c_str(), n, add2.data()));
    StateBuf back(n);
    CALL(wg_state_from_files(n, &st, stor2.data(), snow2.data(), add2.data(), &back.s));
    CALL(wg_unf_write_double(p4.c_str(), n, sb.s.soil));
    std::vector<double> soil_unf(n);
    CALL(wg_unf_read_double(p4.c_str(), n, soil_unf.data()));
    expect(wg_unf_read(p4.c_str(), WG_UNF_F32, int64_t(n) + 5, soil_unf.data()) == WG_E_INPUT, "short file");
    for (const std::string &p : {p1, p2, p3, p4}) std::remove(p.c_str());
    rmdir(dir);
  }
  wgo_destroy(m);
  wg_world_destroy(w);
  std::printf("%s: %d cells, %d days, %d basins, outlet discharge %.6g km3\n", label, n, days, out.n_basins,
              q.empty() ? 0.0 : q[0]);
  return 0;
}

// The calibration search over a synthetic basin whose discharge is linear in gamma.
int search() {
  wg_gamma_search s;
  wg_gamma_search_init(&s);
  float gamma = wg_gamma_start(&s, 2.0f);
  const int years = 5;
  float measured[years], simulated[years];
  for (int y = 0; y < years; ++y) measured[y] = 10.0f + y;
  int calls = 0;
  while (gamma > 0.0f && calls < 40) {
    for (int y = 0; y < years; ++y) simulated[y] = measured[y] * 1.6f / (0.6f + gamma);
    gamma = wg_find_new_gamma(&s, gamma, years, measured, simulated, nullptr, nullptr);
    ++calls;
  }
  expect(calls < 40, "gamma search ends");
  std::printf("gamma search: %d calls, status %d\n", calls, s.calib_status);
  return 0;
}

}  // namespace

int main() {
  int rc = 0;
  rc |= exercise(regional(0, 0), 40, "regional");
  rc |= exercise(regional(5, 1), 40, "regional edge cases");
  rc |= exercise(small_global(), 10, "global 6,000 cells");
  rc |= search();
  if (rc || failures) return 1;
  std::printf("host_sanitize ok\n");
  return 0;
}
