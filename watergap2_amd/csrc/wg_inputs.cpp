// wg_inputs.cpp -- WaterGAP input files on the host: the .UNF grid format, the
// daily climate files, the land-cover tables and the derived soil and
// groundwater statics, so that a WaterGAP input set feeds wg_create and
// wg_load_forcing directly.
//
//   wg_unf_read / wg_unf_write   grid_io_adapters.h:43-129  big-endian raw arrays; double grids
//                                                         are stored as float32
//   wg_read_climate_year         climateYear.cpp:38-54      the four .365.UNF0 files of a year
//   wg_read_lct_tables           daily.cpp:1925-1959, lai.cpp:110-149  LCT_22.DAT, LAI_22.DAT
//   wg_cell_geometry             geo.cpp:7-50               cell area by row, contfreq raised to
//                                                         the surface-water fractions
//   wg_soil_capacity             s_max.cpp:43-74            Smax = TAWC x rooting depth
//   wg_gw_factor                 gw_frac.cpp:31-258         groundwater factor fg and Rgmax
//
// Every formula keeps the reference's element types (float grids, double
// parameters, short/char classes), because the results are the reference's
// grids bit for bit.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/watergap2_amd.h"

namespace {

size_t elem_size(int32_t type) {
  switch (type) {
    case WG_UNF_F32: return 4;
    case WG_UNF_I8: return 1;
    case WG_UNF_I16: return 2;
    case WG_UNF_I32: return 4;
    default: return 0;
  }
}

// byte order of the file (big-endian) <-> host (little-endian x86-64)
void swap_bytes(void *data, size_t size, int64_t count) {
  unsigned char *p = static_cast<unsigned char *>(data);
  if (size == 2) {
    for (int64_t i = 0; i < count; i++) {
      uint16_t v;
      std::memcpy(&v, p + 2 * i, 2);
      v = __builtin_bswap16(v);
      std::memcpy(p + 2 * i, &v, 2);
    }
  } else if (size == 4) {
    for (int64_t i = 0; i < count; i++) {
      uint32_t v;
      std::memcpy(&v, p + 4 * i, 4);
      v = __builtin_bswap32(v);
      std::memcpy(p + 4 * i, &v, 4);
    }
  }
}

int read_raw(const char *path, size_t size, int64_t count, void *out) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return WG_E_INPUT;  // the reference exits: "<file> not found."
  const size_t got = std::fread(out, size, size_t(count), f);
  std::fclose(f);
  // the reference reads whatever the file holds; a file shorter than the grid is
  // an input error here
  if (got != size_t(count)) return WG_E_INPUT;
  swap_bytes(out, size, count);
  return WG_OK;
}

int write_raw(const char *path, size_t size, int64_t count, const void *in) {
  std::vector<unsigned char> buf(size * size_t(count));
  std::memcpy(buf.data(), in, buf.size());
  swap_bytes(buf.data(), size, count);
  FILE *f = std::fopen(path, "wb");
  if (!f) return WG_E_INPUT;
  const size_t put = std::fwrite(buf.data(), 1, buf.size(), f);
  const int closed = std::fclose(f);
  return (put == buf.size() && closed == 0) ? WG_OK : WG_E_INPUT;
}

// reads one whitespace-separated token of a table file
bool next_token(FILE *f, std::string &tok) {
  tok.clear();
  int c;
  while ((c = std::fgetc(f)) != EOF && std::isspace(c)) {
  }
  if (c == EOF) return false;
  do {
    tok.push_back(char(c));
  } while ((c = std::fgetc(f)) != EOF && !std::isspace(c));
  return true;
}

// skips the leading comment lines (those starting with '#')
void skip_comments(FILE *f) {
  int c;
  while ((c = std::fgetc(f)) == '#') {
    while ((c = std::fgetc(f)) != EOF && c != '\n') {
    }
  }
  if (c != EOF) std::ungetc(c, f);
}

template <class T>
bool parse(const std::string &tok, T &v);
template <>
bool parse<double>(const std::string &tok, double &v) {
  char *end = nullptr;
  errno = 0;
  v = std::strtod(tok.c_str(), &end);
  return end && *end == '\0' && errno == 0;
}
template <>
bool parse<float>(const std::string &tok, float &v) {
  char *end = nullptr;
  errno = 0;
  v = std::strtof(tok.c_str(), &end);
  return end && *end == '\0' && errno == 0;
}
template <>
bool parse<int16_t>(const std::string &tok, int16_t &v) {
  char *end = nullptr;
  const long x = std::strtol(tok.c_str(), &end, 10);
  v = int16_t(x);
  return end && *end == '\0' && x >= -32768 && x <= 32767;
}

}  // namespace

extern "C" int wg_unf_type(const char *path) {
  if (!path) return WG_E_ARG;
  const char *dot = std::strrchr(path, '.');
  if (!dot || std::strncmp(dot, ".UNF", 4) != 0 || dot[4] == '\0' || dot[5] != '\0') return WG_E_ARG;
  switch (dot[4]) {
    case '0': return WG_UNF_F32;
    case '1': return WG_UNF_I8;
    case '2': return WG_UNF_I16;
    case '4': return WG_UNF_I32;
    default: return WG_E_ARG;
  }
}

extern "C" int wg_unf_read(const char *path, int32_t type, int64_t count, void *out) {
  const size_t size = elem_size(type);
  if (!path || !out || count < 0 || size == 0) return WG_E_ARG;
  return read_raw(path, size, count, out);
}

extern "C" int wg_unf_write(const char *path, int32_t type, int64_t count, const void *in) {
  const size_t size = elem_size(type);
  if (!path || !in || count < 0 || size == 0) return WG_E_ARG;
  return write_raw(path, size, count, in);
}

extern "C" int wg_unf_read_double(const char *path, int64_t count, double *out) {
  if (!path || !out || count < 0) return WG_E_ARG;
  std::vector<float> buf(static_cast<size_t>(count));
  const int rc = read_raw(path, 4, count, buf.data());
  if (rc) return rc;
  for (int64_t i = 0; i < count; i++) out[i] = static_cast<double>(buf[i]);
  return WG_OK;
}

extern "C" int wg_unf_write_double(const char *path, int64_t count, const double *in) {
  if (!path || !in || count < 0) return WG_E_ARG;
  std::vector<float> buf(static_cast<size_t>(count));
  for (int64_t i = 0; i < count; i++) buf[i] = static_cast<float>(in[i]);
  return write_raw(path, 4, count, buf.data());
}

extern "C" int wg_read_climate_year(const char *climate_dir, int32_t year, int32_t n_cells, float *forcing) {
  if (!climate_dir || n_cells <= 0 || !forcing) return WG_E_ARG;
  // temperature degC, precipitation mm/d, shortwave and longwave W/m2, each a
  // DailyGrid [N][365] (climateYear.cpp:38-54)
  static const char *kName[4] = {"G_TEMP_H08_int_", "G_GPCC_H08day_V20110128_", "G_SSRD_H08_int_",
                                 "G_SLRD_H08_int_"};
  std::vector<float> grid(size_t(n_cells) * 365);
  for (int v = 0; v < 4; v++) {
    const std::string path = std::string(climate_dir) + "/" + kName[v] + std::to_string(year) + ".365.UNF0";
    const int rc = read_raw(path.c_str(), 4, int64_t(n_cells) * 365, grid.data());
    if (rc) return rc;
    for (int32_t d = 0; d < 365; d++)
      for (int32_t i = 0; i < n_cells; i++)
        forcing[(size_t(d) * 4 + v) * n_cells + i] = grid[size_t(i) * 365 + d];
  }
  return WG_OK;
}

extern "C" int wg_read_lct_tables(const char *input_dir, wg_lct_tables *out) {
  if (!input_dir || !out) return WG_E_ARG;
  std::string tok;
  {
    // LCT_22.DAT: class, rooting depth, albedo, snow albedo, degree-day factor,
    // emissivity (daily.cpp:1925-1959)
    const std::string path = std::string(input_dir) + "/LCT_22.DAT";
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return WG_E_INPUT;
    skip_comments(f);
    bool ok = true;
    for (int i = 0; i < WG_NLCT && ok; i++) {
      int16_t idx = 0;
      ok = next_token(f, tok) && parse(tok, idx) && idx - 1 == i;
      ok = ok && next_token(f, tok) && parse(tok, out->rooting_depth[i]);
      ok = ok && next_token(f, tok) && parse(tok, out->albedo[i]);
      ok = ok && next_token(f, tok) && parse(tok, out->albedo_snow[i]);
      ok = ok && next_token(f, tok) && parse(tok, out->ddf[i]);
      ok = ok && next_token(f, tok) && parse(tok, out->emissivity[i]);
    }
    std::fclose(f);
    if (!ok) return WG_E_INPUT;
  }
  {
    // LAI_22.DAT: class, LAI, deciduous fraction, evergreen reduction, initial
    // days, kc min, kc max (lai.cpp:110-149)
    const std::string path = std::string(input_dir) + "/LAI_22.DAT";
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return WG_E_INPUT;
    skip_comments(f);
    bool ok = true;
    for (int i = 0; i < WG_NLCT && ok; i++) {
      int16_t idx = 0;
      ok = next_token(f, tok) && parse(tok, idx) && idx - 1 == i;
      ok = ok && next_token(f, tok) && parse(tok, out->lai[i]);
      ok = ok && next_token(f, tok) && parse(tok, out->decid_frac[i]);
      ok = ok && next_token(f, tok) && parse(tok, out->evergreen_red[i]);
      ok = ok && next_token(f, tok) && parse(tok, out->initial_days[i]);
      ok = ok && next_token(f, tok) && parse(tok, out->kc_min[i]);
      ok = ok && next_token(f, tok) && parse(tok, out->kc_max[i]);
    }
    std::fclose(f);
    if (!ok) return WG_E_INPUT;
  }
  return WG_OK;
}

extern "C" int wg_cell_geometry(int32_t n, const double *area_by_row, const int16_t *row, const double *contfreq_in,
                                const double *loc_lake, const double *glo_lake, const double *loc_res,
                                const double *res_frac, double *cell_area, double *contfreq) {
  if (n < 0 || !area_by_row || !row || !contfreq_in || !cell_area || !contfreq) return WG_E_ARG;
  for (int32_t i = 0; i < n; i++) {
    if (row[i] < 1 || row[i] > 360) return WG_E_INPUT;
    cell_area[i] = area_by_row[row[i] - 1];  // geo.h:33-35
    // geo.cpp:22-44: the surface-water fractions, summed in the reference's order;
    // the continental share is at least their sum
    double water = 0.;
    if (loc_lake) water += loc_lake[i];
    if (glo_lake) water += glo_lake[i];
    if (loc_res) water += loc_res[i];
    if (res_frac) water += res_frac[i];
    contfreq[i] = (water > contfreq_in[i]) ? water : contfreq_in[i];
  }
  return WG_OK;
}

extern "C" int wg_soil_capacity(int32_t n, const float *tawc, const int8_t *land_cover, const double *m_root_d,
                                const float *rooting_depth_lct, float *smax) {
  if (n < 0 || !tawc || !land_cover || !m_root_d || !rooting_depth_lct || !smax) return WG_E_ARG;
  for (int32_t i = 0; i < n; i++) {
    if (land_cover[i] < 1 || land_cover[i] > WG_NLCT) return WG_E_INPUT;
    // rooting depth: parameter multiplier times the class's depth, stored as float
    const float root = float(m_root_d[i] * double(rooting_depth_lct[land_cover[i] - 1]));
    smax[i] = (tawc[i] < 0) ? -9999.f : tawc[i] * root;  // TAWC -9999: no soil data (ice)
  }
  return WG_OK;
}

extern "C" int wg_gw_factor(int32_t n, const int8_t *slope_class, const int8_t *texture, const int8_t *permaglac,
                            const int8_t *aquifer_factor, const float *gw_factor_corr, const double *m_gw_f,
                            const double *m_rg_max, int32_t time_series, float *gw_factor, int16_t *rgmax) {
  if (n < 0 || !slope_class || !texture || !permaglac || !aquifer_factor || !m_gw_f || !m_rg_max || !gw_factor ||
      !rgmax)
    return WG_E_ARG;
  static const short kSlopeClass[7] = {10, 20, 30, 40, 50, 60, 70};
  static const float kSlopeFactor[7] = {1.00f, 0.95f, 0.90f, 0.75f, 0.60f, 0.30f, 0.15f};
  static const short kTexture[3] = {10, 20, 30};
  float rg_table[3] = {5.f, 3.f, 1.5f};
  if (time_series == 0) {  // daily WFD input: the modified Rgmax (gw_frac.cpp:111-115)
    rg_table[0] = 7.f;
    rg_table[1] = 4.5f;
    rg_table[2] = 2.5f;
  }
  static const float kTextureFactor[3] = {1.f, 0.95f, 0.70f};
  for (int32_t i = 0; i < n; i++) {
    // slope factor, interpolated between the classes (gw_frac.cpp:40-76); classes
    // outside 10..70 keep 0 (the reference reports them and goes on)
    const signed char sc = (slope_class[i] == 0) ? 10 : slope_class[i];
    float slope_factor = 0.f;
    if (kSlopeClass[6] == sc) {
      slope_factor = kSlopeFactor[6];
    } else {
      for (int k = 0; k <= 5; k++) {
        if (kSlopeClass[k] == sc) {
          slope_factor = kSlopeFactor[k];
          break;
        } else if (sc > kSlopeClass[k] && sc < kSlopeClass[k + 1]) {
          slope_factor = kSlopeFactor[k] + (((kSlopeFactor[k + 1] - kSlopeFactor[k]) /
                                             (kSlopeClass[k + 1] - kSlopeClass[k])) * (sc - kSlopeClass[k]));
          break;
        }
      }
    }
    // texture factor and Rgmax (gw_frac.cpp:88-150)
    const double m_rg = m_rg_max[i];
    float texture_factor = 0.f, rg_f = 0.f;
    const signed char tx = texture[i];
    if (tx <= 0 || tx == 2) {
      texture_factor = 0.95f;
      rg_f = 3;
    } else if (tx == 1) {
      texture_factor = 0;
      rg_f = 0;
    } else {
      for (int k = 0; k <= 2; k++) {
        if (kTexture[k] == tx) {
          texture_factor = kTextureFactor[k];
          rg_f = float(m_rg * rg_table[k]);
          break;
        } else if (k < 2 && tx > kTexture[k] && tx < kTexture[k + 1]) {
          texture_factor = kTextureFactor[k] + (((kTextureFactor[k + 1] - kTextureFactor[k]) /
                                                 (kTexture[k + 1] - kTexture[k])) * (tx - kTexture[k]));
          rg_f = float((m_rg * rg_table[k]) +
                       ((m_rg * (rg_table[k + 1] - rg_table[k]) / (kTexture[k + 1] - kTexture[k])) * (tx - kTexture[k])));
          break;
        }
      }
    }
    // groundwater factor (gw_frac.cpp:160-200)
    if (texture_factor < 0) {
      gw_factor[i] = -99;
    } else {
      float sf = slope_factor, tf = texture_factor;
      float af = float((short)aquifer_factor[i] / 100.0);
      float pf = float(1. - ((float)permaglac[i] / 100.));
      if (sf < 0) sf = 0;
      if (sf > 1) sf = 1;
      if (af < 0) af = 0;
      if (af > 1) af = 1;
      if (tf < 0) tf = 0;
      if (tf > 1) tf = 1;
      if (pf < 0) pf = 0;
      if (pf > 1) pf = 1;
      gw_factor[i] = float(m_gw_f[i] * sf * tf * af * pf);
      if (gw_factor[i] > 1.) gw_factor[i] = 0.95f;
    }
    // Rgmax in mm/d x 100, rounded half up (gw_frac.cpp:19-22, 202-207)
    rgmax[i] = (rg_f >= 0) ? (int16_t)std::floor(rg_f * 100 + 0.5) : int16_t(-9999);
    // the correction grid (gw_frac.cpp:224-233)
    if (gw_factor_corr && gw_factor_corr[i] > 0.) {
      gw_factor[i] = float(m_gw_f[i] * gw_factor_corr[i]);
      if (gw_factor[i] > 1.) gw_factor[i] = 0.95f;
    }
  }
  return WG_OK;
}
