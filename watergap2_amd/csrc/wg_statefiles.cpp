// wg_statefiles.cpp -- the model state in the reference's restart files (text
// variants; the NetCDF-4 variants need the NetCDF/HDF5 libraries, which this
// build does not use):
//
//   WGHM water storage states   wghmStateFile.cpp:156-229 (load), 357-413 (saveMean)
//   snow in elevation           snowInElevationFile.cpp:80-123 (load), 180-210 (save)
//   additional out- and input   additionalOutputInputFile.cpp:186-240 (load), 310-400 (save)
//
// and the conversion between a member's state (wg_state) and the files' values:
// the storages are kept in the files in mm over the cell's continental area
// (routing.cpp:4998-5020, integrateWGHM.cpp:830-846) and read back by
// dailyWaterBalanceClass::setStorages (daily.cpp:1896-1924) and
// routingClass::setStorages (routing.cpp:851-882); the LAI state and the release
// coefficient travel in the additional file (lai.cpp:55-64, 165-173,
// routing.cpp:388-393, 3080).
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/watergap2_amd.h"

namespace {

constexpr int kWidth = 33;
const char *kStateNames[11] = {"TWS", "CANOPY", "SNOW", "SOIL", "LOCALLAKE", "LOCALWETLAND", "GLOBALLAKE",
                               "GLOBALWETLAND", "RESERVOIR", "RIVER", "GROUNDWATER"};
// column labels of the additional file, as the reference writes them
// (additionalOutputInputFile.cpp:328-381)
const char *kAddNames[53] = {
    "G_days_since_start", "G_GrowingStatus", "G_PrecSum", "G_totalUnsatisfiedUse", "G_UnsatisfiedUsePrevYear",
    "K_release", "G_landAreaFrac", "G_landAreaFracPrevTimestep", "G_locWetlAreaReductionFactor",
    "G_gloLakeEvapoReductionFactor", "G_groundwaterStorage", "G_locLakeAreaReductionFactor",
    "G_gloWetlAreaReductionFactor", "G_gloResEvapoReductionFactor", "G_fswbInit[n]", "G_gloWetlStorage",
    "G_fswbLandAreaFrac", "G_locWetlStorage[n]", "G_locLakeStorage[n]", "G_riverStorage[n]",
    "G_soilwatercontent[n]", "G_gloLakeStorage[n]", "G_canopywatercontent[n]", "G_gloResStorage[n]",
    "G_snow[n]", "G_withdrawalIrrigFromSwb[n]", "G_consumptiveUseIrrigFromSwb[n]", "G_unsatAllocUSE",
    "G_AllocUSE", "G_secondcell", "G_daily_UnsatAllocUseNextDay[n]", " G_daily_allocatedUseNextDay[n]",
    " G_AllocUSETOneigborcell[n]", " G_fswbLandAreaFracNextTimestep[n]", " G_PrevUnsatAllocUse[n]",
    " G_dailyRemainingUse", " G_dailyAllocatedUse[n]", " G_PrevTotalUnstatisfieduse[n]",
    " G_reducedReturnFlow[n]", " G_unsatisfiedNAsFromIrrig[n]", " G_dailySatisAllocatedUseInSecondCell",
    " G_dailyAllocatedUse", " G_unsatUseRiparian[n]", " G_ActualUse[secondCell]", " G_glores_prevyear[n]",
    " G_fLocLake[n]", " G_fGloWet[n]", " G_fLocWet[n]", " G_unsatUseRiparian[i]",
    " G_unsatisfiedNAsFromOtherSectors[n]", " G_reducedReturnFlowPrevYear[n]",
    " G_unsatisfiedNAsFromIrrigPrevYear[n]", " G_unsatisfiedNAsFromOtherSectorsPrevYear[n]"};

// std::setw(w) << std::scientific with precision 16 = printf("%*.16e")
void put_sci(std::string &s, double v, int width) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%*.16e", width, v);
  s += buf;
}

void put_right(std::string &s, const char *text, int width) {
  char buf[128];
  std::snprintf(buf, sizeof(buf), "%*s", width, text);
  s += buf;
}

int write_text(const char *path, const std::string &s) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return WG_E_INPUT;
  const size_t put = std::fwrite(s.data(), 1, s.size(), f);
  const int closed = std::fclose(f);
  return (put == s.size() && closed == 0) ? WG_OK : WG_E_INPUT;
}

// The data rows of a state file: lines whose first non-blank character is not a
// letter, each an id followed by `cols` numbers; the reference stops at the first
// empty line.  Rows go to row `id - 1` when by_id, else in file order.
int read_rows(const char *path, int32_t n, int cols, bool by_id, double *out) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return WG_E_INPUT;
  std::string line;
  int32_t row = 0;
  int rc = WG_OK;
  std::vector<char> buf(1 << 16);
  while (rc == WG_OK) {
    line.clear();
    bool got = false;
    while (std::fgets(buf.data(), int(buf.size()), f)) {
      got = true;
      line += buf.data();
      if (!line.empty() && line.back() == '\n') break;
    }
    if (!got) break;
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
    if (line.empty()) break;
    const char *p = line.c_str();
    while (*p && std::isspace(static_cast<unsigned char>(*p))) p++;
    if (std::isalpha(static_cast<unsigned char>(*p))) continue;  // header lines
    char *end = nullptr;
    const long id = std::strtol(p, &end, 10);
    if (end == p) {
      rc = WG_E_INPUT;
      break;
    }
    const int32_t r = by_id ? int32_t(id - 1) : row;
    if (r < 0 || r >= n) {
      rc = WG_E_INPUT;
      break;
    }
    p = end;
    for (int k = 0; k < cols; k++) {
      const double v = std::strtod(p, &end);
      if (end == p) {
        rc = WG_E_INPUT;
        break;
      }
      out[int64_t(r) * cols + k] = v;
      p = end;
    }
    row++;
  }
  std::fclose(f);
  if (rc == WG_OK && row != n) rc = WG_E_INPUT;
  return rc;
}

}  // namespace

extern "C" int wg_write_wghm_state_txt(const char *path, int32_t n, const double *storages_mm) {
  if (!path || n < 0 || !storages_mm) return WG_E_ARG;
  std::string s = "WGHM water storage states\n";
  put_right(s, "ID", 6);
  for (const char *name : kStateNames) put_right(s, name, kWidth);
  s += "\n";
  for (int32_t i = 0; i < n; i++) {
    const double *v = storages_mm + int64_t(i) * WG_NSTORE;
    char id[32];
    std::snprintf(id, sizeof(id), "%6d", i + 1);
    s += id;
    // Cell::tws (wghmStateFile.cpp:730-735): the sum in this order
    const double tws = v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7] + v[8] + v[9];
    put_sci(s, tws, kWidth);
    for (int k = 0; k < WG_NSTORE; k++) put_sci(s, v[k], kWidth);
    s += "\n";
  }
  return write_text(path, s);
}

extern "C" int wg_read_wghm_state_txt(const char *path, int32_t n, double *storages_mm) {
  if (!path || n < 0 || !storages_mm) return WG_E_ARG;
  // rows of (TWS, 10 storages), placed by their cell id
  std::vector<double> rows(size_t(n) * 11);
  const int rc = read_rows(path, n, 11, true, rows.data());
  if (rc) return rc;
  for (int32_t i = 0; i < n; i++)
    for (int k = 0; k < WG_NSTORE; k++) storages_mm[int64_t(i) * WG_NSTORE + k] = rows[size_t(i) * 11 + 1 + k];
  return WG_OK;
}

extern "C" int wg_write_snow_elevation_txt(const char *path, int32_t n, const double *snow) {
  if (!path || n < 0 || !snow) return WG_E_ARG;
  std::string s = "Snow in elevation for land cells\n";
  char buf[64];
  for (int32_t i = 0; i < n; i++) {
    std::snprintf(buf, sizeof(buf), "%d\t", i + 1);
    s += buf;
    for (int j = 0; j < 100; j++) {
      std::snprintf(buf, sizeof(buf), "%.16e\t", snow[int64_t(i) * WG_NELEV + j]);
      s += buf;
    }
    std::snprintf(buf, sizeof(buf), "%.16e", snow[int64_t(i) * WG_NELEV + 100]);
    s += buf;
    if (i < n - 1) s += "\n";
  }
  return write_text(path, s);
}

extern "C" int wg_read_snow_elevation_txt(const char *path, int32_t n, double *snow) {
  if (!path || n < 0 || !snow) return WG_E_ARG;
  // rows in file order (snowInElevationFile.cpp:101-121 counts them)
  return read_rows(path, n, WG_NELEV, false, snow);
}

extern "C" int wg_write_additional_txt(const char *path, int32_t n, const double *cols) {
  if (!path || n < 0 || !cols) return WG_E_ARG;
  std::string s = "additional out- and input\n";
  put_right(s, "ID", 6);
  for (const char *name : kAddNames) put_right(s, name, kWidth);
  s += "\n";
  char id[32];
  for (int32_t i = 0; i < n; i++) {
    std::snprintf(id, sizeof(id), "%6d", i + 1);
    s += id;
    for (int j = 0; j < WG_NADDITIONAL; j++) put_sci(s, cols[int64_t(i) * WG_NADDITIONAL + j], kWidth);
    if (i < n - 1) s += "\n";
  }
  return write_text(path, s);
}

extern "C" int wg_read_additional_txt(const char *path, int32_t n, double *cols) {
  if (!path || n < 0 || !cols) return WG_E_ARG;
  return read_rows(path, n, WG_NADDITIONAL, false, cols);
}

extern "C" int wg_state_to_files(int32_t n, const wg_static_fields *st, const wg_state *s, double *storages_mm,
                                 double *snow_mm, double *additional) {
  if (n < 0 || !st || !s || !st->cell_area || !st->contfreq) return WG_E_ARG;
  for (int32_t i = 0; i < n; i++) {
    const double cf = st->contfreq[i], laf = s->laf ? s->laf[i] : 0.;
    // km3 -> mm over the continental area (routing.cpp:5005-5019)
    const double mm = (st->cell_area[i] * (cf / 100.)) / 1000000.;
    if (storages_mm) {
      double *v = storages_mm + int64_t(i) * WG_NSTORE;
      // land storages per land area -> per continental area (integrateWGHM.cpp:843-845)
      v[0] = s->canopy ? s->canopy[i] * laf / cf : 0.;
      v[1] = s->snow ? s->snow[i] * laf / cf : 0.;
      v[2] = s->soil ? s->soil[i] * laf / cf : 0.;
      v[3] = s->loc_lake ? s->loc_lake[i] / mm : 0.;
      v[4] = s->loc_wetl ? s->loc_wetl[i] / mm : 0.;
      v[5] = s->glo_lake ? s->glo_lake[i] / mm : 0.;
      v[6] = s->glo_wetl ? s->glo_wetl[i] / mm : 0.;
      v[7] = s->reservoir ? s->reservoir[i] / mm : 0.;
      v[8] = s->river ? s->river[i] / mm : 0.;
      v[9] = s->groundwater ? s->groundwater[i] / mm : 0.;
    }
    if (snow_mm && s->snow_elev) {
      // integrateWGHM.cpp:830-841
      for (int e = 0; e < WG_NELEV; e++)
        snow_mm[int64_t(i) * WG_NELEV + e] = (laf <= 0.) ? 0. : s->snow_elev[int64_t(i) * WG_NELEV + e] * laf / cf;
    }
    if (additional) {
      double *a = additional + int64_t(i) * WG_NADDITIONAL;
      for (int j = 0; j < WG_NADDITIONAL; j++) a[j] = 0.;
      if (s->lai_days) a[0] = s->lai_days[i];
      if (s->lai_status) a[1] = s->lai_status[i];
      if (s->lai_precsum) a[2] = s->lai_precsum[i];
      a[5] = s->k_release ? s->k_release[i] : 0.1;
      a[6] = laf;
      if (s->laf_prev) a[7] = s->laf_prev[i];
      if (s->red_loc_wetl) a[8] = s->red_loc_wetl[i];
      if (s->red_glo_lake) a[9] = s->red_glo_lake[i];
      if (s->groundwater) a[10] = s->groundwater[i];
      if (s->red_loc_lake) a[11] = s->red_loc_lake[i];
      if (s->red_glo_wetl) a[12] = s->red_glo_wetl[i];
      if (s->red_res) a[13] = s->red_res[i];
      // routing.cpp:5215-5222: storages and the SWB fraction
      a[14] = (st->loc_lake[i] / 100.) + (st->loc_wetland[i] / 100.) + (st->glo_wetland[i] / 100.);
      if (s->glo_wetl) a[15] = s->glo_wetl[i];
      if (s->loc_wetl) a[17] = s->loc_wetl[i];
      if (s->loc_lake) a[18] = s->loc_lake[i];
      if (s->river) a[19] = s->river[i];
      if (s->soil) a[20] = s->soil[i];
      if (s->glo_lake) a[21] = s->glo_lake[i];
      if (s->canopy) a[22] = s->canopy[i];
      if (s->reservoir) a[23] = s->reservoir[i];
      if (s->snow) a[24] = s->snow[i];
      if (s->fswb_frac_next) a[33] = s->fswb_frac_next[i];
      a[36] = st->glo_lake[i] / 100.;
    }
  }
  return WG_OK;
}

extern "C" int wg_state_from_files(int32_t n, const wg_static_fields *st, const double *storages_mm,
                                   const double *snow_mm, const double *additional, wg_state *s) {
  if (n < 0 || !st || !s || !st->cell_area || !st->contfreq) return WG_E_ARG;
  for (int32_t i = 0; i < n; i++) {
    const double *a = additional ? additional + int64_t(i) * WG_NADDITIONAL : nullptr;
    // the land-area fraction the storages were saved with
    const double laf = a ? a[6] : (s->laf ? s->laf[i] : 0.);
    const double cf = st->contfreq[i];
    const double mm = (st->cell_area[i] * (cf / 100.)) / 1000000.;
    if (a) {
      if (s->lai_days) s->lai_days[i] = int32_t(a[0]);
      if (s->lai_status) s->lai_status[i] = int32_t(a[1]);
      if (s->lai_precsum) s->lai_precsum[i] = a[2];
      if (s->k_release) s->k_release[i] = a[5];
      if (s->laf) s->laf[i] = a[6];
      if (s->laf_prev) s->laf_prev[i] = a[7];
      if (s->red_loc_wetl) s->red_loc_wetl[i] = a[8];
      if (s->red_glo_lake) s->red_glo_lake[i] = a[9];
      if (s->red_loc_lake) s->red_loc_lake[i] = a[11];
      if (s->red_glo_wetl) s->red_glo_wetl[i] = a[12];
      if (s->red_res) s->red_res[i] = a[13];
      if (s->fswb_frac_next) s->fswb_frac_next[i] = a[33];
    }
    if (storages_mm) {
      const double *v = storages_mm + int64_t(i) * WG_NSTORE;
      // daily.cpp:1907-1921
      if (s->canopy) s->canopy[i] = (laf <= 0.) ? 0. : v[0] * cf / laf;
      if (s->snow) s->snow[i] = (laf <= 0.) ? 0. : v[1] * cf / laf;
      if (s->soil) s->soil[i] = (laf <= 0.) ? 0. : v[2] * cf / laf;
      // routing.cpp:858-867
      if (s->loc_lake) s->loc_lake[i] = v[3] * mm;
      if (s->loc_wetl) s->loc_wetl[i] = v[4] * mm;
      if (s->glo_lake) s->glo_lake[i] = v[5] * mm;
      if (s->glo_wetl) s->glo_wetl[i] = v[6] * mm;
      if (s->reservoir) s->reservoir[i] = v[7] * mm;
      if (s->river) s->river[i] = v[8] * mm;
      if (s->groundwater) s->groundwater[i] = v[9] * mm;
    }
    if (snow_mm && s->snow_elev)
      for (int e = 0; e < WG_NELEV; e++)
        s->snow_elev[int64_t(i) * WG_NELEV + e] = (laf <= 0.) ? 0. : snow_mm[int64_t(i) * WG_NELEV + e] * cf / laf;
  }
  return WG_OK;
}
