// abi_driver.cpp -- a compiled C++ caller of the library's C ABI.
//
// Built with g++ against include/watergap2_amd.h (and the world generator's
// header) and linked to libwatergap2_amd.so only, the way a C++ host such as the
// reference's integrateWGHM.cpp would bind it.
//
//   abi_driver check              host-only calls (no device is touched)
//   abi_driver run <dir> <days>   a 12 x 15 regional grid (seed 0) for <days> days
//                                 on the GPU; writes the member-0 state and the
//                                 station discharge to <dir> as raw arrays
//
// Exit status 0 on success; every failing call prints its name, code and message.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <unistd.h>

#include "watergap2_amd.h"
#include "../watergap2_amd/csrc/wg_world.h"

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

// The grid of the Python fixture regional_spec(seed=0, nrows=12, ncols=15,
// glolake_every=10, res_every=20).
wg_world_spec small_regional() {
  wg_world_spec s;
  std::memset(&s, 0, sizeof s);
  s.kind = 0;
  s.nrows = 12;
  s.ncols = 15;
  s.row0 = 81;
  s.col0 = 361;
  s.seed = 0;
  s.water_frac_max = 0.1;
  s.glolake_every = 10;
  s.res_every = 20;
  return s;
}

template <class T>
bool write_raw(const std::string &path, const T *p, size_t count) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  size_t w = std::fwrite(p, sizeof(T), count, f);
  return std::fclose(f) == 0 && w == count;
}

int check() {
  wg_options_profile opt;
  wg_default_options(&opt);
  expect(opt.time_series == 1 && opt.pet_opt == 0 && opt.river_velo_opt == 2 && opt.res_opt == 1 &&
             opt.subtract_use == 0 && opt.fractional_routing_opt == 1,
         "wg_default_options gives the parity profile");

  // routing schedule of a generated grid, recomputed through the ABI
  wg_world_spec spec = small_regional();
  wg_world *w = nullptr;
  CALL(wg_world_create(&spec, &w));
  const int32_t n = w->n;
  std::vector<int32_t> inflow9(size_t(n) * 9), outflow(n), order(n), level(n);
  CALL(wg_inflow_outflow(n, w->ldd, w->row, w->col, w->cell_at, inflow9.data(), outflow.data()));
  CALL(wg_rout_order(n, inflow9.data(), outflow.data(), order.data(), level.data()));
  expect(std::memcmp(outflow.data(), w->downstream, sizeof(int32_t) * n) == 0, "outflow cells");
  expect(std::memcmp(order.data(), w->rout_order, sizeof(int32_t) * n) == 0, "routing order");
  for (int32_t i = 0; i < n; ++i)
    if (outflow[i] > 0) expect(order[outflow[i] - 1] > order[i], "downstream cell comes later");

  // a cycle is refused
  std::vector<int32_t> cyc_in(18, 0), cyc_out = {2, 1};
  cyc_in[0] = 2;
  cyc_in[9] = 1;
  std::vector<int32_t> cyc_order(2);
  expect(wg_rout_order(2, cyc_in.data(), cyc_out.data(), cyc_order.data(), nullptr) == WG_E_CYCLE,
         "two-cell cycle gives WG_E_CYCLE");

  // .UNF round trip of a big-endian int16 grid and a float32-stored double grid
  char tmpl[] = "/tmp/wg_abi_driver_XXXXXX";
  char *dir = mkdtemp(tmpl);
  expect(dir != nullptr, "temporary directory");
  if (dir) {
    std::string p2 = std::string(dir) + "/G_TEST.UNF2";
    std::string p0 = std::string(dir) + "/G_TEST.UNF0";
    expect(wg_unf_type(p2.c_str()) == WG_UNF_I16, "type of .UNF2");
    CALL(wg_unf_write(p2.c_str(), WG_UNF_I16, n, w->row));
    std::vector<int16_t> row_back(n);
    CALL(wg_unf_read(p2.c_str(), WG_UNF_I16, n, row_back.data()));
    expect(std::memcmp(row_back.data(), w->row, sizeof(int16_t) * n) == 0, ".UNF2 round trip");
    FILE *f = std::fopen(p2.c_str(), "rb");
    unsigned char b[2] = {0, 0};
    expect(f && std::fread(b, 1, 2, f) == 2, "read back .UNF2 bytes");
    if (f) std::fclose(f);
    expect(b[0] == ((uint16_t(w->row[0]) >> 8) & 0xff) && b[1] == (uint16_t(w->row[0]) & 0xff),
           ".UNF2 is big-endian");
    CALL(wg_unf_write_double(p0.c_str(), n, w->cell_area));
    std::vector<double> area_back(n);
    CALL(wg_unf_read_double(p0.c_str(), n, area_back.data()));
    for (int32_t i = 0; i < n; ++i)
      expect(area_back[i] == double(float(w->cell_area[i])), ".UNF0 double grid is float32 on disk");
    expect(wg_unf_read(p0.c_str(), WG_UNF_F32, n + 1, area_back.data()) == WG_E_INPUT,
           "short .UNF0 file gives WG_E_INPUT");
    std::remove(p2.c_str());
    std::remove(p0.c_str());
    rmdir(dir);
  }

  // host-side calibration and surrogate entry points
  wg_gamma_search gs;
  wg_gamma_search_init(&gs);
  expect(gs.gamma_upper_limit == 5.0f && gs.gamma_lower_limit == 0.1f, "gamma search limits");
  expect(wg_gamma_start(&gs, 2.0f) > 0.0f, "gamma start");
  expect(wg_mlp_n_params() > 0, "surrogate parameter count");

  wg_world_destroy(w);
  if (failures) return 1;
  std::printf("abi_driver check ok (%d cells)\n", n);
  return 0;
}

int run(const std::string &dir, int days) {
  wg_world_spec spec = small_regional();
  wg_world *w = nullptr;
  CALL(wg_world_create(&spec, &w));
  const int32_t n = w->n;
  wg_static_fields st;
  std::memset(&st, 0, sizeof st);
  wg_world_statics(w, &st);
  wg_options_profile opt;
  wg_default_options(&opt);

  wg_handle *h = nullptr;
  int rc = wg_create(n, &st, &opt, 1, &h);
  if (rc != 0) {
    std::fprintf(stderr, "FAIL: wg_create returned %d: %s\n", rc, h ? wg_last_error(h) : "");
    return 1;
  }
  CALL(wg_set_stations(h, w->n_basins, w->basin_outlet));
  std::vector<float> forcing(size_t(days) * 4 * n);
  CALL(wg_world_forcing(w, 0, 0, days, forcing.data()));

  // invalid forcing is refused and leaves the loaded days alone
  std::vector<float> bad(forcing.begin(), forcing.begin() + 4 * size_t(n));
  bad[size_t(n) + 3] = std::nanf("");
  expect(wg_load_forcing(h, bad.data(), 1) == WG_E_INPUT, "NaN forcing gives WG_E_INPUT");
  bad[size_t(n) + 3] = -1.0f;
  expect(wg_load_forcing(h, bad.data(), 1) == WG_E_INPUT, "negative precipitation gives WG_E_INPUT");

  CALL(wg_load_forcing(h, forcing.data(), days));
  CALL(wg_step_days(h, 0, 1, days));
  CALL(wg_synchronize(h));

  std::vector<double> d[25];
  for (auto &v : d) v.assign(n, 0.0);
  std::vector<double> snow_elev(size_t(n) * 101);
  std::vector<int32_t> lai_days(n), lai_status(n);
  wg_state s;
  std::memset(&s, 0, sizeof s);
  double **dbl[] = {&s.canopy, &s.soil, &s.snow, &s.storage_transfer, &s.lai_precsum, &s.groundwater,
                    &s.loc_lake, &s.loc_wetl, &s.glo_lake, &s.glo_wetl, &s.reservoir, &s.river,
                    &s.red_loc_lake, &s.red_loc_wetl, &s.red_glo_lake, &s.red_glo_wetl, &s.red_res,
                    &s.red_river, &s.river_frac_next, &s.fswb_frac_next, &s.laf, &s.laf_prev,
                    &s.k_release};
  const char *names[] = {"canopy", "soil", "snow", "storage_transfer", "lai_precsum", "groundwater",
                         "loc_lake", "loc_wetl", "glo_lake", "glo_wetl", "reservoir", "river",
                         "red_loc_lake", "red_loc_wetl", "red_glo_lake", "red_glo_wetl", "red_res",
                         "red_river", "river_frac_next", "fswb_frac_next", "laf", "laf_prev",
                         "k_release"};
  const int n_dbl = int(sizeof(names) / sizeof(names[0]));
  for (int k = 0; k < n_dbl; ++k) *dbl[k] = d[k].data();
  s.snow_elev = snow_elev.data();
  s.lai_days = lai_days.data();
  s.lai_status = lai_status.data();
  CALL(wg_get_state(h, 0, &s));

  std::vector<double> sum(w->n_basins);
  CALL(wg_get_station_discharge(h, sum.data(), nullptr, 0));

  bool ok = true;
  for (int k = 0; k < n_dbl; ++k) ok &= write_raw(dir + "/" + names[k] + ".bin", d[k].data(), size_t(n));
  ok &= write_raw(dir + "/snow_elev.bin", snow_elev.data(), snow_elev.size());
  ok &= write_raw(dir + "/lai_days.bin", lai_days.data(), size_t(n));
  ok &= write_raw(dir + "/lai_status.bin", lai_status.data(), size_t(n));
  ok &= write_raw(dir + "/started.bin", &s.started, 1);
  ok &= write_raw(dir + "/discharge.bin", sum.data(), sum.size());
  expect(ok, "writing the outputs");

  const int32_t n_stations = w->n_basins;
  CALL(wg_destroy(h));
  wg_world_destroy(w);
  if (failures) return 1;
  std::printf("abi_driver run ok (%d cells, %d days, %d stations)\n", n, days, n_stations);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc >= 2 && std::strcmp(argv[1], "check") == 0) return check();
  if (argc >= 4 && std::strcmp(argv[1], "run") == 0) return run(argv[2], std::atoi(argv[3]));
  std::fprintf(stderr, "usage: abi_driver check | abi_driver run <dir> <days>\n");
  return 2;
}
