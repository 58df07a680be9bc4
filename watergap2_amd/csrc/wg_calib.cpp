// wg_calib.cpp -- the per-basin runoff-coefficient search of the calibration loop.
//
// Restates calibGammaClass::findNewGamma (calibration.cpp:266-636) with its
// function-static state moved into a per-basin struct, so that many basins can be
// searched side by side.  Arithmetic is in float as in the reference (sums,
// averages, bracket midpoints); every decision and update is kept.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <vector>

#include "../../include/watergap2_amd.h"

extern "C" void wg_gamma_search_init(wg_gamma_search *s) {
  if (!s) return;
  std::memset(s, 0, sizeof(*s));
  // calibration.cpp:275-284 (function statics), calibration.h:13-15, 48, calibration.cpp:45-48
  s->gamma_low = -99.f;
  s->gamma_high = -99.f;
  s->sum_diff_old = 0.f;
  s->call_counter = 0;
  s->gamma_cond = 0;
  s->calib_status = 0;
  s->cell_corr_factor_ind = 1;
  s->cell_corr_factor = 1.f;
  s->gamma_upper_limit = 5.f;
  s->gamma_lower_limit = 0.1f;
  s->last_nash_sutcliffe = 0.f;
  s->last_sum_diff = 0.f;
  s->measured_avg_adapt = 0.f;
  s->sim_runoff_avg = 0.f;
}

// calibration.cpp:196-205: start value of gamma from the parameter file
extern "C" float wg_gamma_start(const wg_gamma_search *s, float gamma_from_file) {
  if (gamma_from_file < 0) return 1.f;
  if ((gamma_from_file > s->gamma_upper_limit) || (gamma_from_file < s->gamma_lower_limit)) return 1.f;
  return gamma_from_file;
}

extern "C" float wg_find_new_gamma(wg_gamma_search *s, float gamma_old, int32_t n_years, const float *measured,
                                   const float *simulated, const float *sim_inflow, const float *sim_water_use) {
  if (!s || n_years <= 0 || !measured || !simulated) return -99.f;
  const float up_lim = s->gamma_upper_limit, lo_lim = s->gamma_lower_limit;
  float &gamma_low = s->gamma_low, &gamma_high = s->gamma_high;
  s->call_counter++;
  const int callCounter = s->call_counter;

  // calibration.cpp:300-315: sums over the years with observations
  float sumOfDifferences = 0, measuredRunoffSum = 0, simInflowSum = 0, simWaterUseSum = 0, simRunoffSum = 0;
  int numberOfYears = 0;
  for (int i = 0; i < n_years; i++) {
    if (measured[i] > -1) {
      numberOfYears++;
      sumOfDifferences += simulated[i] - measured[i];
      measuredRunoffSum += measured[i];
      simRunoffSum += simulated[i];
      simInflowSum += sim_inflow ? sim_inflow[i] : 0.f;
      simWaterUseSum += sim_water_use ? sim_water_use[i] : 0.f;
    }
  }
  const float measuredRunoffAvg = measuredRunoffSum / numberOfYears;
  // the 1 % criterion of calibration.cpp:327, 378, 459
  const bool within_1pc = (std::fabs(sumOfDifferences / (numberOfYears * measuredRunoffAvg))) < 0.01;

  // The reference reads the still unset local `gamma` in the second call
  // (calibration.cpp:409, 437); any value but the limit itself takes the
  // midpoint, which is what a fresh local holds in practice.
  float gamma = -98.f;
  if (1 == callCounter) {
    // calibration.cpp:318-366
    if (within_1pc) {
      gamma = -99;
      s->gamma_cond = 1;
      s->calib_status = 1;
    } else if (sumOfDifferences > 0) {
      if (gamma_old >= up_lim) {
        gamma = -99;
      } else {
        gamma_low = gamma_old;
        gamma = up_lim;
      }
    } else {
      if (gamma_old <= lo_lim) {
        gamma = -99;
      } else {
        gamma_high = gamma_old;
        gamma = lo_lim;
      }
    }
  } else if (2 == callCounter) {
    // calibration.cpp:367-448
    if (within_1pc) {
      gamma = -99;
      s->gamma_cond = 1;
      s->calib_status = 1;
    } else if (sumOfDifferences > 0) {
      if (gamma_old == up_lim) {
        gamma = gamma_old;
        gamma_high = up_lim;
      } else if (gamma_old < up_lim) {
        if (gamma_high < 0) {
          gamma_low = gamma_old;
          gamma = gamma_old * 2.0;
        } else {
          gamma_low = gamma_old;
          if (gamma != up_lim) gamma = (gamma_low + gamma_high) / 2.0;
        }
      }
    } else {
      if (gamma_old == lo_lim) {
        gamma = gamma_old;
        gamma_low = lo_lim;
      } else if (gamma_old > lo_lim) {
        if (gamma_low < 0) {
          gamma = gamma_old / 2.0;
          gamma_high = gamma_old;
        } else {
          gamma_high = gamma_old;
          if (gamma != lo_lim) gamma = (gamma_low + gamma_high) / 2.0;
        }
      }
    }
  } else {
    // calibration.cpp:449-523
    if (within_1pc) {
      gamma = -99;
      s->gamma_cond = 1;
      s->calib_status = 1;
    } else if (sumOfDifferences > 0) {
      if (gamma_old == up_lim) {
        gamma = -99;
        s->gamma_cond = 2;
      } else if (gamma_old < up_lim) {
        if (gamma_high < 0) {
          gamma_low = gamma_old;
          gamma = gamma_old * 2.0;
        } else {
          gamma_low = gamma_old;
          gamma = (gamma_low + gamma_high) / 2.0;
        }
      }
    } else {
      if (gamma_old == lo_lim) {
        gamma = -99;
        s->gamma_cond = 2;
      } else if (gamma_old > lo_lim) {
        if (gamma_low < 0) {
          gamma = gamma_old / 2.0;
          gamma_high = gamma_old;
        } else {
          gamma_high = gamma_old;
          gamma = (gamma_low + gamma_high) / 2.0;
        }
      }
    }
  }

  // calibration.cpp:527-544: Nash-Sutcliffe efficiency (more than one year)
  float NashSutclCoeff;
  if (n_years > 1) {
    float sum1 = 0, sum2 = 0;
    for (int i = 0; i < n_years; i++) {
      if (measured[i] > -1) {
        sum1 += (measured[i] - measuredRunoffAvg) * (measured[i] - measuredRunoffAvg);
        sum2 += (simulated[i] - measured[i]) * (simulated[i] - measured[i]);
      }
    }
    NashSutclCoeff = (sum1 - sum2) / sum1;
  } else {
    NashSutclCoeff = -99;
  }

  // calibration.cpp:549-597: at a limit, the 10 % rule and the cell correction factor
  float measuredRunoffAvgAdapt = measuredRunoffAvg;
  if ((gamma_old == up_lim) && (s->gamma_cond == 2)) {
    measuredRunoffAvgAdapt = measuredRunoffAvg * 1.1;
    if ((simRunoffSum / numberOfYears) < measuredRunoffAvgAdapt)
      s->calib_status = 2;
    else
      s->cell_corr_factor_ind = 99;
  }
  if ((gamma_old == lo_lim) && (s->gamma_cond == 2)) {
    measuredRunoffAvgAdapt = measuredRunoffAvg * 0.9;
    if ((simRunoffSum / numberOfYears) > measuredRunoffAvgAdapt)
      s->calib_status = 2;
    else
      s->cell_corr_factor_ind = 99;
  }
  if (s->cell_corr_factor_ind == 99) {
    s->cell_corr_factor = (measuredRunoffAvgAdapt + (simWaterUseSum - simInflowSum) / numberOfYears) /
                          ((simRunoffSum + simWaterUseSum - simInflowSum) / numberOfYears);
  }

  // calibration.cpp:613-632: termination rules
  if ((gamma > 0) && (gamma < (0.99 * lo_lim)) && (callCounter > 2)) gamma = -99;
  if ((gamma > 0) && (std::fabs((s->sum_diff_old / sumOfDifferences) - 1) < 0.0001)) gamma = -99;
  s->sum_diff_old = sumOfDifferences;
  if ((gamma < 0) && (std::fabs(s->cell_corr_factor - 1) > 0.01)) s->calib_status = 3;

  s->last_nash_sutcliffe = NashSutclCoeff;
  s->last_sum_diff = sumOfDifferences;
  s->measured_avg_adapt = measuredRunoffAvgAdapt;
  s->sim_runoff_avg = simRunoffSum / numberOfYears;  // createCorrectionGrid's argument (calibration.cpp:627)
  return gamma;
}

// ---- calibration basins and upstream stations ----


namespace {

// upstreamBasin (calib_basins.cpp:302-378): claims for basin `b` the cells upstream
// of station cell `cn` (1-based) that hold the station cell's previous value (0, or
// the basin the station lies in, which the new basin then splits).  Returns the
// overlap code: 1 when the station lay in an earlier basin, 2 when earlier basins
// lie upstream, 3 both, 0 neither.  Each cell has one downstream cell, so the set
// of claimed cells does not depend on the order in which the reference's sweeps
// visit them.
int claim_upstream(int32_t n, const int32_t *inflow9, int16_t *basin, int32_t cn, int16_t b) {
  const int16_t base = basin[cn - 1];
  const bool upstream = base != 0;
  bool downstream = false;
  std::vector<int32_t> stack{cn - 1}, claimed;
  basin[cn - 1] = -1;
  while (!stack.empty()) {
    const int32_t c = stack.back();
    stack.pop_back();
    claimed.push_back(c);
    for (int j = 0; j < 9; j++) {
      const int32_t u = inflow9[size_t(c) * 9 + j];
      if (u <= 0 || u > n) continue;
      if (basin[u - 1] == base) {
        basin[u - 1] = -1;
        stack.push_back(u - 1);
      } else {
        downstream = true;
      }
    }
  }
  for (int32_t c : claimed) basin[c] = b;
  return (upstream ? 1 : 0) + (downstream ? 2 : 0);
}

}  // namespace

extern "C" int wg_calib_basins(int32_t n, const int32_t *inflow9, const int32_t *outflow, int32_t n_stations,
                               const int32_t *station_cells, int32_t *used, int16_t *calib_basin, int32_t *overlap) {
  if (n <= 0 || !inflow9 || !outflow || n_stations < 0 || (n_stations > 0 && !station_cells) || !calib_basin)
    return WG_E_ARG;
  if (n_stations > 32767) return WG_E_ARG;  // basin numbers are int16 (G_CALIB_BASIN.UNF2)
  std::fill(calib_basin, calib_basin + n, int16_t(0));
  std::vector<int32_t> accepted;
  for (int32_t k = 0; k < n_stations; k++) {
    const int32_t cn = station_cells[k];
    if (used) used[k] = 0;
    if (overlap) overlap[k] = -1;
    // calib_basins.cpp:117-131: a cell outside the land mask (cell number <= 1) and a
    // cell that is already a basin's station are skipped
    if (cn <= 1 || cn > n) continue;
    bool seen = false;
    for (int32_t a : accepted) seen = seen || (a == cn);
    if (seen) continue;
    accepted.push_back(cn);
    const int code = claim_upstream(n, inflow9, calib_basin, cn, int16_t(accepted.size()));
    if (used) used[k] = 1;
    if (overlap) overlap[k] = code;
  }
  // calib_basins.cpp:270-285: the cell downstream of each station that belongs to no
  // basin is marked -1 (its inflow is computed too)
  for (int32_t cn : accepted) {
    const int32_t o = outflow[cn - 1];
    if (o > 0 && o <= n && calib_basin[o - 1] == 0) calib_basin[o - 1] = -1;
  }
  return int(accepted.size());
}

extern "C" int wg_upstream_stations(int32_t n, const int32_t *inflow9, int32_t n_basins, const int32_t *basin_cells,
                                    int32_t all_stations, int32_t *offsets, int32_t *stations, int32_t max_entries) {
  if (n <= 0 || !inflow9 || n_basins < 0 || (n_basins > 0 && !basin_cells) || !offsets || max_entries < 0 ||
      (max_entries > 0 && !stations))
    return WG_E_ARG;
  for (int32_t b = 0; b < n_basins; b++)
    if (basin_cells[b] < 1 || basin_cells[b] > n) return WG_E_ARG;
  // upstreamStationClass::findStations (upstream_stations.cpp:66-140).  The reference
  // sweeps all cells in index order, processing cells marked 1 (marking their inflow
  // cells 1, or, for a station cell, all_stations), until a sweep leaves none; a cell
  // marked during a sweep is processed in the same sweep when its index lies ahead.
  // Here a sweep takes its pending cells from an ordered set in the same order.  The
  // marks persist from one basin to the next, as in the reference.  Each basin's list
  // holds the stations in reverse order of discovery (new entries go to the front).
  std::vector<int8_t> done(size_t(n), 0);
  std::vector<int32_t> first_basin(size_t(n), 0);  // 1-based basin of a station cell (first in order)
  for (int32_t b = n_basins - 1; b >= 0; b--) first_basin[basin_cells[b] - 1] = b + 1;
  int32_t total = 0;
  offsets[0] = 0;
  for (int32_t cb = 0; cb < n_basins; cb++) {
    std::vector<int32_t> found;
    std::set<int32_t> cur, next;
    done[basin_cells[cb] - 1] = 1;
    for (int32_t c = 0; c < n; c++)
      if (done[c] == 1) cur.insert(c);
    while (!cur.empty()) {
      while (!cur.empty()) {
        const int32_t c = *cur.begin();
        cur.erase(cur.begin());
        if (done[c] != 1) continue;
        done[c] = 2;
        for (int i = 0; i < 9; i++) {
          const int32_t u = inflow9[size_t(c) * 9 + i];
          if (u == 0) continue;
          if (u < 0 || u > n) return WG_E_INPUT;
          const int32_t st = first_basin[u - 1];
          const int8_t mark = st ? int8_t(all_stations ? 1 : 0) : int8_t(1);
          done[u - 1] = mark;
          if (st) found.push_back(st);
          if (mark == 1) {
            if (u - 1 > c) {
              cur.insert(u - 1);
            } else {
              next.insert(u - 1);
            }
          }
        }
      }
      // the cells left marked 1 for the next sweep (a cell queued there may have
      // been processed since)
      for (int32_t c : next)
        if (done[c] == 1) cur.insert(c);
      next.clear();
    }
    for (auto it = found.rbegin(); it != found.rend(); ++it) {
      if (total >= max_entries) return WG_E_ARG;
      stations[total++] = *it;
    }
    offsets[cb + 1] = total;
  }
  return WG_OK;
}

// calibGammaClass::writeCorrFactors (calibration.cpp:656-728): the station correction
// factor after a search, in float as the reference computes it.
// calibGammaClass::createCorrectionGrid (calibration.cpp:730-801), in float as the
// reference computes it: the mean of the annual grids (read back from their float
// files), the basin's sum of absolute means, and per basin cell
// 1 - sign(mean) * (simulated - measured) / sum, limited to 0.5 .. 1.5.
extern "C" int wg_cell_corr_factor_grid(int32_t n_cells, const int16_t *calib_basin, int32_t station_number,
                                        int32_t n_years, const float *annual_pot, float simulated, float measured,
                                        float *mean_pot, double *corr) {
  if (n_cells <= 0 || n_years <= 0 || n_years > 32767 || !calib_basin || !annual_pot || !corr) return WG_E_ARG;
  const short numberOfYears = short(n_years);
  std::vector<float> mean(size_t(n_cells), 0.0f);
  for (int32_t y = 0; y < n_years; y++)
    for (int32_t i = 0; i < n_cells; i++) mean[i] += annual_pot[size_t(y) * n_cells + i];
  for (int32_t i = 0; i < n_cells; i++) mean[i] /= numberOfYears;
  float sumOfAbsValuesForBasin = 0;
  for (int32_t i = 0; i < n_cells; i++)
    if (station_number == calib_basin[i]) sumOfAbsValuesForBasin += std::fabs(mean[i]);
  for (int32_t i = 0; i < n_cells; i++) {
    if (station_number != calib_basin[i]) continue;
    const int sg = mean[i] > 0 ? 1 : (mean[i] == 0 ? 0 : -1);  // sign(), calibration.cpp:32-43
    corr[i] = 1 - ((sg * (simulated - measured)) / sumOfAbsValuesForBasin);
    if (corr[i] > 1.5) corr[i] = 1.5;
    if (corr[i] < 0.5) corr[i] = 0.5;
  }
  if (mean_pot) std::copy(mean.begin(), mean.end(), mean_pot);
  return WG_OK;
}

extern "C" float wg_station_corr_factor(wg_gamma_search *s, float gamma, int32_t n_years, const float *measured,
                                        const float *simulated) {
  if (!s || n_years <= 0 || !measured || !simulated) return -99.f;
  float measured_sum = 0, sim_sum = 0;
  for (int32_t i = 0; i < n_years; i++) {
    if (measured[i] > -1) {
      measured_sum += measured[i];
      sim_sum += simulated[i];
    }
  }
  const bool cfa_done = s->cell_corr_factor_ind == 99;
  if ((gamma == s->gamma_upper_limit) && cfa_done) measured_sum = measured_sum * 1.1;
  if ((gamma == s->gamma_lower_limit) && cfa_done) measured_sum = measured_sum * 0.9;
  float cfs = cfa_done ? measured_sum / sim_sum : 1.0f;
  if (((cfs > 1.0) && (cfs < 1.01)) || ((cfs < 1.0) && (cfs > 0.99))) cfs = 1.0;
  if ((cfs > 1.0) || (cfs < 1.0)) s->calib_status = 4;
  return cfs;
}
