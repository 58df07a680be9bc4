// wg_prepare.cpp -- routing preprocessing on the host: the steps of
// prepare_routing_files (rout_prepare.cpp:61-336) that turn a flow-direction map
// into the routing inputs of the day loop.
//
//   wg_ldd_from_flowdir     rout_prepare.cpp:68-112   Arc/Info flow codes -> WaterGAP LDD
//   wg_replace_flow_to_ocean rout_prepare.cpp:287-336 flow into the ocean -> sink
//   wg_flow_accumulation    rout_prepare.cpp:441-494  upstream cells + 1 (0 inside loops)
//   wg_derive_basins        rout_prepare.cpp:497-571  basin of every cell, cells to the outlet
//   wg_reindex_basins       rout_prepare.cpp:573-604  basins of more than one cell renumbered 1..
//   wg_neighbour_cells      rout_prepare.cpp:672-692  the 8 neighbours of every cell
//   wg_cell_distances       rout_prepare.cpp:694-752  distances to the neighbours by row
//   wg_river_slope          rout_prepare.cpp:754-788  slope along the flow direction
//   wg_river_length         rout_prepare.cpp:790-832  river length with meandering
//   wg_alloc_coeff          rout_prepare.cpp:895-944  reservoirs' shares of downstream demand
//   wg_reservoir_start_month rout_prepare.cpp:950-1019 start of the operational year
//   wg_prepare_routing      rout_prepare.cpp:61-336   the whole pipeline
//
// The reference computes the sweeps of flow accumulation and basin derivation by
// rescanning every cell per sweep (O(N * depth)); here they run over frontiers
// (O(N)), visiting cells in the order that gives the reference's values.  The
// float/double mix of the distance, slope and length formulas is the
// reference's, because the results are stored as float grids and compared bit
// for bit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/watergap2_amd.h"

namespace {

inline int32_t lattice_cell(const int32_t *cell_at, int col, int row) {
  if (row < 1 || row > 360) return 0;  // GCRC_2 border rows (rout_prepare.cpp:132-146)
  if (col < 1) col += 720;            // longitude wrap
  if (col > 720) col -= 720;
  return cell_at[(row - 1) * 720 + (col - 1)];
}

}  // namespace

extern "C" int wg_ldd_from_flowdir(int32_t n, const int16_t *flowdir, int8_t *ldd) {
  if (n < 0 || !flowdir || !ldd) return WG_E_ARG;
  for (int32_t i = 0; i < n; i++) {
    int8_t v;
    switch (flowdir[i]) {
      case -1: v = -1; break;  // inland sink
      case 8: v = 1; break;    // SW
      case 4: v = 2; break;    // S
      case 2: v = 3; break;    // SE
      case 16: v = 4; break;   // W
      case 0: v = 5; break;    // sink
      case 1: v = 6; break;    // E
      case 32: v = 7; break;   // NW
      case 64: v = 8; break;   // N
      case 128: v = 9; break;  // NE
      default: v = 5; break;   // no data: a sink
    }
    ldd[i] = v;
  }
  return WG_OK;
}

extern "C" int wg_replace_flow_to_ocean(int32_t n, const int16_t *row, const int16_t *col, const int32_t *cell_at,
                                        int8_t *ldd) {
  if (n < 0 || !row || !col || !cell_at || !ldd) return WG_E_ARG;
  // a cell whose flow direction points at a lattice position without land (or
  // past the poles) becomes a sink (rout_prepare.cpp:287-336)
  static const int kD[10][2] = {{0, 0}, {-1, 1}, {0, 1}, {1, 1}, {-1, 0}, {0, 0}, {1, 0}, {-1, -1}, {0, -1}, {1, -1}};
  for (int32_t i = 0; i < n; i++) {
    const int c = ldd[i];
    if (c == 5 || c == -1 || c < 1 || c > 9) continue;
    if (lattice_cell(cell_at, col[i] + kD[c][0], row[i] + kD[c][1]) == 0) ldd[i] = 5;
  }
  return WG_OK;
}

extern "C" int wg_flow_accumulation(int32_t n, const int32_t *inflow9, int16_t *flow_acc) {
  if (n < 0 || !inflow9 || !flow_acc) return WG_E_ARG;
  // A cell gets its value once every listed inflow cell has one: 1 plus their
  // values, summed in slot order in 16-bit arithmetic as the reference's short
  // grid does.  Cells on a loop, or fed by one, never get a value and stay 0.
  std::vector<int32_t> pending(n, 0), head(n, -1), next_edge, edge_to;
  for (int32_t i = 0; i < n; i++)
    for (int k = 0; k < 9; k++) {
      const int32_t u = inflow9[int64_t(i) * 9 + k];
      if (u < 0 || u > n) return WG_E_ARG;
      if (u == 0) continue;
      pending[i]++;
      // edge u -> i, so that finishing u releases i
      edge_to.push_back(i);
      next_edge.push_back(head[u - 1]);
      head[u - 1] = int32_t(edge_to.size()) - 1;
    }
  std::vector<int32_t> ready;
  for (int32_t i = 0; i < n; i++) {
    flow_acc[i] = 0;
    if (pending[i] == 0) ready.push_back(i);
  }
  while (!ready.empty()) {
    const int32_t i = ready.back();
    ready.pop_back();
    int16_t acc = 1;
    for (int k = 0; k < 9; k++) {
      const int32_t u = inflow9[int64_t(i) * 9 + k];
      if (u != 0) acc = int16_t(acc + flow_acc[u - 1]);
    }
    // the reference marks "no value yet" with 0; a sum that wraps to 0 reads as
    // unfinished there too, so its receivers wait for ever
    flow_acc[i] = acc;
    if (acc == 0) continue;
    for (int32_t e = head[i]; e >= 0; e = next_edge[e])
      if (--pending[edge_to[e]] == 0) ready.push_back(edge_to[e]);
  }
  return WG_OK;
}

extern "C" int wg_derive_basins(int32_t n, int8_t *ldd, const int32_t *inflow9, uint16_t *basin, uint16_t *dist,
                                int32_t *n_basins) {
  if (n < 0 || !ldd || !inflow9 || !basin || !n_basins) return WG_E_ARG;
  // done: -99 untouched, 1 on the current front, 2 finished (rout_prepare.cpp:513-571)
  std::vector<int8_t> done(n, -99);
  int32_t nb = 0;
  for (int32_t i = 0; i < n; i++) {
    basin[i] = 0;
    if (dist) dist[i] = 0;
    if (ldd[i] == 0 || ldd[i] == -99) {
      // cells outside the LDD map become sinks, and so endpoints below
      done[i] = 2;
      ldd[i] = 5;
    }
    if (ldd[i] == 5 || ldd[i] == -1) {
      done[i] = 1;
      nb++;
      basin[i] = uint16_t(nb);
    }
  }
  if (nb > 60000 - 1) return WG_E_INPUT;  // nb_max (rout_prepare.cpp:25, 540-543)
  std::vector<int32_t> front, marked;
  for (int32_t i = 0; i < n; i++)
    if (done[i] == 1) front.push_back(i);
  int32_t step = 1;
  std::vector<int8_t> in_marked(n, 0);
  while (!front.empty()) {
    // the front in ascending cell order, as the reference's scan meets it
    marked.clear();
    for (int32_t i : front) {
      if (done[i] != 1) continue;
      if (dist) dist[i] = uint16_t(step - 1);
      done[i] = 2;
      for (int k = 0; k < 9; k++) {
        const int32_t u = inflow9[int64_t(i) * 9 + k];
        if (u < 0 || u > n) return WG_E_ARG;
        if (u == 0) continue;
        basin[u - 1] = basin[i];
        done[u - 1] = -1;
        if (!in_marked[u - 1]) {
          in_marked[u - 1] = 1;
          marked.push_back(u - 1);
        }
      }
    }
    front.clear();
    for (int32_t u : marked) {
      in_marked[u] = 0;
      if (done[u] == -1) {
        done[u] = 1;
        front.push_back(u);
      }
    }
    std::sort(front.begin(), front.end());
    step++;
  }
  *n_basins = nb;
  return WG_OK;
}

extern "C" int wg_reindex_basins(int32_t n, int32_t n_basins, uint16_t *basin, int32_t *n_multi) {
  if (n < 0 || n_basins < 0 || n_basins >= 60000 || !basin || !n_multi) return WG_E_ARG;
  // cells per basin in the reference's 16-bit counters (rout_prepare.cpp:202-208)
  std::vector<int16_t> cells(60000, 0);
  std::vector<int32_t> first(n_basins + 1, -1);
  for (int32_t i = 0; i < n; i++) {
    cells[basin[i]] = int16_t(cells[basin[i]] + 1);
    if (basin[i] <= n_basins && first[basin[i]] < 0) first[basin[i]] = i;
  }
  std::vector<int32_t> renum(n_basins + 1, -1);
  int32_t j = 0;
  for (int32_t b = 1; b <= n_basins; b++)
    if (cells[b] != 1) renum[b] = ++j;
  if (j > 10000 - 1) return WG_E_INPUT;  // nb2_max (rout_prepare.cpp:26, 598-601)
  // a basin counted as one cell needs no routing: its (first) cell gets basin 0
  for (int32_t b = 1; b <= n_basins; b++)
    if (cells[b] == 1 && first[b] >= 0) basin[first[b]] = 0;
  for (int32_t i = 0; i < n; i++)
    if (basin[i] >= 1 && basin[i] <= n_basins && renum[basin[i]] > 0) basin[i] = uint16_t(renum[basin[i]]);
  *n_multi = j;
  return WG_OK;
}

extern "C" int wg_neighbour_cells(int32_t n, const int16_t *row, const int16_t *col, const int32_t *cell_at,
                                  int32_t *neighbour8) {
  if (n < 0 || !row || !col || !cell_at || !neighbour8) return WG_E_ARG;
  static const int kD[8][2] = {{1, 0}, {1, -1}, {0, -1}, {-1, -1}, {-1, 0}, {-1, 1}, {0, 1}, {1, 1}};
  for (int32_t i = 0; i < n; i++)
    for (int k = 0; k < 8; k++)
      neighbour8[int64_t(i) * 8 + k] = lattice_cell(cell_at, col[i] + kD[k][0], row[i] + kD[k][1]);
  return WG_OK;
}

extern "C" int wg_cell_distances(float *dist) {
  if (!dist) return WG_E_ARG;
  // dist[k * 360 + r]: distance (km) from a cell of row r + 1 to its neighbour in
  // LDD direction k + 1, on a sphere of the Earth's volume (GCELLDIST.9.UNF0)
  auto D = [&](int k, int r) -> float & { return dist[k * 360 + r]; };
  const float pi = 3.141592653589793f;
  const float earth_radius = 6371.211f;
  const float vert_dist = float(2 * pi * earth_radius * 0.5 / 360.0);
  for (int i = 0; i <= 359; i++) {
    const float l = float(0.25 + i * 0.5);
    const float horiz_dist = float(2 * pi * earth_radius * std::sin(l * pi / 180.0) * 0.5 / 360.0);
    D(1, i) = vert_dist;
    D(3, i) = horiz_dist;
    D(4, i) = 0;
    D(5, i) = horiz_dist;
    D(7, i) = vert_dist;
  }
  // diagonals from the horizontal distances of the two rows and the vertical one,
  // in float (the reference's sqrt of a float argument)
  for (int i = 0; i <= 358; i++) D(0, i) = std::sqrt(D(3, i) * D(3, i + 1) + D(1, i) * D(1, i));
  D(0, 359) = -99;
  for (int i = 1; i <= 359; i++) D(6, i) = std::sqrt(D(3, i) * D(3, i - 1) + D(1, i) * D(1, i));
  D(6, 0) = -99;
  for (int i = 0; i <= 359; i++) {
    D(2, i) = D(0, i);
    D(8, i) = D(6, i);
  }
  return WG_OK;
}

extern "C" int wg_river_slope(int32_t n, const float *meander, const float *dist, const float *altitude,
                              const int32_t *outflow, const int8_t *ldd, const int16_t *row, float *slope) {
  if (n < 0 || !meander || !dist || !altitude || !outflow || !ldd || !row || !slope) return WG_E_ARG;
  for (int32_t i = 0; i < n; i++) {
    const int32_t o = outflow[i];
    if (o < 0 || o > n) return WG_E_ARG;
    if (o != 0 && ldd[i] != 5 && ldd[i] != -1) {
      if (ldd[i] < 1 || ldd[i] > 9 || row[i] < 1 || row[i] > 360) return WG_E_INPUT;
      const float drop = altitude[i] - altitude[o - 1];
      const float mean_meander = (meander[i] + meander[o - 1]) / 2;
      slope[i] = float(drop / (1000.0 * dist[(ldd[i] - 1) * 360 + (row[i] - 1)] * mean_meander));
    } else {
      slope[i] = 0;
    }
  }
  const float minslope = 0.00001f;
  for (int32_t i = 0; i < n; i++)
    if (slope[i] < minslope) slope[i] = minslope;
  return WG_OK;
}

extern "C" int wg_river_length(int32_t n, const float *meander, const int16_t *row, const int16_t *col,
                               const float *dist, const int32_t *outflow, float *length) {
  if (n < 0 || !meander || !row || !col || !dist || !outflow || !length) return WG_E_ARG;
  auto D = [&](int k, int r) { return dist[k * 360 + r]; };
  for (int32_t i = 0; i < n; i++) {
    const int32_t o = outflow[i] - 1;
    if (o < -1 || o >= n || row[i] < 1 || row[i] > 360) return WG_E_ARG;
    float len;
    if (o == -1)
      len = 55.f;  // outlets and inland sinks
    else if (row[i] == row[o])
      len = D(3, row[i] - 1);  // east or west
    else if (row[i] > row[o])
      len = (col[i] == col[o]) ? D(7, row[i] - 1) : D(8, row[i] - 1);  // north, diagonal north
    else
      len = (col[i] == col[o]) ? D(1, row[i] - 1) : D(2, row[i] - 1);  // south, diagonal south
    if (o == -1)
      len *= meander[i];
    else
      len *= ((meander[i] + meander[o]) / 2);
    length[i] = len;
  }
  return WG_OK;
}

extern "C" int wg_alloc_coeff(int32_t n, const int32_t *outflow, const double *reservoir_area,
                              const double *mean_outflow, double *alloc_coeff) {
  if (n < 0 || !outflow || !reservoir_area || !mean_outflow || !alloc_coeff) return WG_E_ARG;
  std::vector<double> inflow_res(n, 0.);
  for (int64_t k = 0; k < int64_t(n) * 5; k++) alloc_coeff[k] = 1.0;
  for (int32_t i = 0; i < n; i++)
    if (outflow[i] < 0 || outflow[i] > n) return WG_E_ARG;
  // the mean outflow of every reservoir reaches up to 5 cells downstream, up to
  // the next reservoir; each such cell sums what reaches it, in cell order
  for (int32_t i = 0; i < n; i++) {
    if (!(reservoir_area[i] > 0)) continue;
    int k = 0;
    int32_t d = outflow[i];
    while (k < 5 && d > 0 && reservoir_area[d - 1] <= 0) {
      inflow_res[d - 1] += mean_outflow[i];
      d = outflow[d - 1];
      k++;
    }
  }
  // a reservoir's share of each of those cells' demand
  for (int32_t i = 0; i < n; i++) {
    if (!(reservoir_area[i] > 0)) continue;
    int k = 0;
    int32_t d = outflow[i];
    while (k < 5 && d > 0 && reservoir_area[d - 1] <= 0) {
      alloc_coeff[int64_t(i) * 5 + k] = mean_outflow[i] / inflow_res[d - 1];
      d = outflow[d - 1];
      k++;
    }
  }
  return WG_OK;
}

extern "C" int wg_reservoir_start_month(int32_t n, const double *mean_outflow_monthly, const double *mean_outflow,
                                        int8_t *start_month) {
  if (n < 0 || !mean_outflow_monthly || !mean_outflow || !start_month) return WG_E_ARG;
  // "begin of the longest dry period" (variante 3): months whose mean is below
  // the annual mean are dry; the scan starts at the first wet month so that a run
  // across the turn of the year stays whole.  As in the reference, the counter of
  // dry months is reset only when a run beats the longest so far (a shorter run
  // adds its months to the next one), and a tie keeps the earlier run.
  for (int32_t i = 0; i < n; i++) {
    const double *q = mean_outflow_monthly + int64_t(i) * 12;
    const double annual = mean_outflow[i];
    int start = 0, dry_length = 0, start_ = 0, counter_length = 0;
    bool last_dry = true;
    int month = 0;
    while (month < 12 && q[month] < annual) month++;
    const int beg_month = month;
    for (int m = 0; m < 12; m++) {
      month = m + beg_month;
      if (month > 11) month -= 12;
      if (q[month] < annual) {
        counter_length++;
        if (!last_dry) {
          start_ = month;
          last_dry = true;
        }
      } else if (last_dry) {
        last_dry = false;
        if (counter_length > dry_length) {
          start = start_;
          dry_length = counter_length;
          counter_length = 0;
        }
      }
    }
    if (counter_length > dry_length) start = start_;
    // (the reference leaves `start` unset when no month is dry; 0 here: January)
    start_month[i] = int8_t(start + 1);
  }
  return WG_OK;
}

extern "C" int wg_prepare_routing(const wg_routing_inputs *in, wg_routing_files *out) {
  if (!in || !out || in->n_cells <= 0 || !in->row || !in->col || !in->cell_at || !in->altitude || !in->meander ||
      (!in->flowdir && !in->ldd))
    return WG_E_ARG;
  if (!out->ldd || !out->inflow9 || !out->outflow || !out->flow_acc || !out->basins || !out->rout_order)
    return WG_E_ARG;
  const int32_t n = in->n_cells;
  int rc;
  // rout_prepare.cpp:68-115: flow directions
  if (in->flowdir) {
    rc = wg_ldd_from_flowdir(n, in->flowdir, out->ldd);
    if (rc) return rc;
  } else {
    for (int32_t i = 0; i < n; i++) out->ldd[i] = in->ldd[i];
  }
  // rout_prepare.cpp:155-178: inflow cells and flow accumulation; cells on a loop
  // (no accumulation) become sinks, and both are computed again
  std::vector<int32_t> tmp_out(n);
  rc = wg_inflow_outflow(n, out->ldd, in->row, in->col, in->cell_at, out->inflow9, tmp_out.data());
  if (rc) return rc;
  rc = wg_flow_accumulation(n, out->inflow9, out->flow_acc);
  if (rc) return rc;
  bool corrected = false;
  for (int32_t i = 0; i < n; i++)
    if (out->flow_acc[i] == 0) {
      out->ldd[i] = 5;
      out->flow_acc[i] = 1;
      corrected = true;
    }
  if (corrected) {
    rc = wg_inflow_outflow(n, out->ldd, in->row, in->col, in->cell_at, out->inflow9, tmp_out.data());
    if (rc) return rc;
    rc = wg_flow_accumulation(n, out->inflow9, out->flow_acc);
    if (rc) return rc;
  }
  // rout_prepare.cpp:180-197: basins, renumbered to the basins of several cells
  std::vector<uint16_t> dist(n, 0);
  int32_t nb = 0, nb2 = 0;
  rc = wg_derive_basins(n, out->ldd, out->inflow9, out->basins, out->cells_to_outlet ? out->cells_to_outlet : dist.data(),
                        &nb);
  if (rc) return rc;
  rc = wg_reindex_basins(n, nb, out->basins, &nb2);
  if (rc) return rc;
  out->n_basins = nb2;
  // rout_prepare.cpp:244-261: outflow cells of the final LDD, distances, slope, length
  std::vector<int32_t> in9_final(size_t(n) * 9);
  rc = wg_inflow_outflow(n, out->ldd, in->row, in->col, in->cell_at, in9_final.data(), out->outflow);
  if (rc) return rc;
  std::vector<float> cdist(9 * 360);
  float *dist9 = out->cell_dist ? out->cell_dist : cdist.data();
  rc = wg_cell_distances(dist9);
  if (rc) return rc;
  if (out->river_slope) {
    rc = wg_river_slope(n, in->meander, dist9, in->altitude, out->outflow, out->ldd, in->row, out->river_slope);
    if (rc) return rc;
  }
  if (out->river_length) {
    rc = wg_river_length(n, in->meander, in->row, in->col, dist9, out->outflow, out->river_length);
    if (rc) return rc;
  }
  // rout_prepare.cpp:263: routing order from the inflow cells of the accumulation
  // step and the final outflow cells
  rc = wg_rout_order(n, out->inflow9, out->outflow, out->rout_order, nullptr);
  if (rc) return rc;
  // rout_prepare.cpp:265-267, 887-1022: reservoirs
  if (in->reservoir_area && in->mean_outflow) {
    if (out->alloc_coeff) {
      rc = wg_alloc_coeff(n, out->outflow, in->reservoir_area, in->mean_outflow, out->alloc_coeff);
      if (rc) return rc;
    }
    if (out->start_month && in->mean_outflow_monthly) {
      rc = wg_reservoir_start_month(n, in->mean_outflow_monthly, in->mean_outflow, out->start_month);
      if (rc) return rc;
    }
  }
  return WG_OK;
}

// routingClass::annualInit, Parts 1-2 (routing.cpp:1015-1100): the reservoirs and
// regulated lakes in operation in the year to use.
extern "C" int wg_reservoir_year(int32_t n, int32_t year_to_use, const double *area_full, const double *cap_full,
                                 const int32_t *start_year, const int8_t *reg_lake, double *area, double *cap) {
  if (n <= 0 || !area_full || !cap_full || !start_year || !area || !cap) return WG_E_ARG;
  for (int32_t i = 0; i < n; i++) {
    area[i] = 0.;  // routing.cpp:1015-1016
    cap[i] = 0.;
    if (reg_lake && reg_lake[i] == 1) {  // routing.cpp:1040-1048
      area[i] = area_full[i];
      cap[i] = cap_full[i];
    }
    if (year_to_use >= start_year[i]) {  // routing.cpp:1091-1095
      area[i] = area_full[i];
      cap[i] = cap_full[i];
    }
  }
  return WG_OK;
}
