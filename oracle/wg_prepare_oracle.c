/*
 * wg_prepare_oracle.c -- CPU oracle of the routing preprocessing (test
 * infrastructure only; see wg_oracle.h).  A plain restatement of the reference's
 * sweeps in prepare_routing_files (rout_prepare.cpp), each function citing the
 * lines it follows, with the reference's own scan-every-cell-per-sweep loops.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "wg_oracle.h"

/* rout_prepare.cpp:441-494 */
int wgo_flow_accumulation(int32_t n, const int32_t *inflow9, int16_t *acc) {
  int32_t cells_left = 99999, cells_left_prev;
  for (int32_t i = 0; i < n; i++) acc[i] = 1;
  for (int32_t i = 0; i < n; i++)
    for (int k = 0; k <= 8; k++)
      if (inflow9[(int64_t)i * 9 + k] != 0) acc[i] = 0;
  do {
    cells_left_prev = cells_left;
    cells_left = 0;
    for (int32_t i = 0; i < n; i++) {
      if (acc[i] != 0) continue;
      int later = 0;
      for (int k = 0; k <= 8; k++) {
        int32_t u = inflow9[(int64_t)i * 9 + k];
        if (u != 0 && acc[u - 1] == 0) later = 1;
      }
      if (!later) {
        short fa = 1;
        for (int k = 0; k <= 8; k++) {
          int32_t u = inflow9[(int64_t)i * 9 + k];
          if (u != 0) fa = (short)(fa + acc[u - 1]);
        }
        acc[i] = fa;
      } else {
        cells_left++;
      }
    }
  } while (cells_left > 0 && cells_left_prev - cells_left != 0);
  return 0;
}

/* rout_prepare.cpp:497-571 */
int wgo_derive_basins(int32_t n, int8_t *ldd, const int32_t *inflow9, uint16_t *ws, uint16_t *dist,
                      int32_t *n_basins) {
  signed char *done = (signed char *)malloc((size_t)(n > 0 ? n : 1));
  int32_t i = 0, k;
  for (int32_t c = 0; c < n; c++) {
    done[c] = -99;
    ws[c] = 0;
    dist[c] = 0;
  }
  for (int32_t c = 0; c < n; c++) {
    if (ldd[c] == 0 || ldd[c] == -99) {
      done[c] = 2;
      ws[c] = 0;
      ldd[c] = 5;
    }
    if (ldd[c] == 5 || ldd[c] == -1) {
      done[c] = 1;
      i++;
      ws[c] = (uint16_t)i;
    }
  }
  *n_basins = i;
  i = 1;
  do {
    k = 0;
    for (int32_t c = 0; c < n; c++) {
      if (done[c] == 1) {
        dist[c] = (uint16_t)(i - 1);
        k++;
        done[c] = 2;
        for (int j = 0; j <= 8; j++) {
          int32_t u = inflow9[(int64_t)c * 9 + j];
          if (u != 0) {
            ws[u - 1] = ws[c];
            done[u - 1] = -1;
          }
        }
      }
    }
    for (int32_t c = 0; c < n; c++)
      if (done[c] == -1) done[c] = 1;
    i++;
  } while (k != 0);
  free(done);
  return 0;
}

/* rout_prepare.cpp:573-604 (and the 16-bit cell counts of 202-208) */
int wgo_reindex_basins(int32_t n, int32_t nb, uint16_t *ws, int32_t *n_multi) {
  short *cells = (short *)calloc(60000, sizeof(short));
  int32_t j = 0;
  for (int32_t c = 0; c < n; c++) cells[ws[c]]++;
  for (int32_t b = 1; b <= nb; b++) {
    if (cells[b] == 1) {
      for (int32_t c = 0; c < n; c++)
        if (ws[c] == b) {
          ws[c] = 0;
          break;
        }
    } else {
      j++;
      for (int32_t c = 0; c < n; c++)
        if (ws[c] == b) ws[c] = (uint16_t)j;
    }
  }
  free(cells);
  *n_multi = j;
  return 0;
}

/* rout_prepare.cpp:694-752; dist[k*360 + r] */
void wgo_cell_distances(float *dist) {
  const float pi = 3.141592653589793f;
  const float earth_radius = 6371.211f;
  float vert_dist, horiz_dist, l;
  vert_dist = 2 * pi * earth_radius * 0.5 / 360.0;
  for (int i = 0; i <= 359; i++) {
    l = 0.25 + i * 0.5;
    horiz_dist = 2 * pi * earth_radius * sin(l * pi / 180.0) * 0.5 / 360.0;
    dist[1 * 360 + i] = vert_dist;
    dist[3 * 360 + i] = horiz_dist;
    dist[4 * 360 + i] = 0;
    dist[5 * 360 + i] = horiz_dist;
    dist[7 * 360 + i] = vert_dist;
  }
  for (int i = 0; i <= 358; i++)
    dist[0 * 360 + i] = sqrtf(dist[3 * 360 + i] * dist[3 * 360 + i + 1] + dist[1 * 360 + i] * dist[1 * 360 + i]);
  dist[0 * 360 + 359] = -99;
  for (int i = 1; i <= 359; i++)
    dist[6 * 360 + i] = sqrtf(dist[3 * 360 + i] * dist[3 * 360 + i - 1] + dist[1 * 360 + i] * dist[1 * 360 + i]);
  dist[6 * 360 + 0] = -99;
  for (int i = 0; i <= 359; i++) {
    dist[2 * 360 + i] = dist[0 * 360 + i];
    dist[8 * 360 + i] = dist[6 * 360 + i];
  }
}

/* rout_prepare.cpp:754-788 */
void wgo_river_slope(int32_t n, const float *meander, const float *dist, const float *alt, const int32_t *outflow,
                     const int8_t *ldd, const int16_t *row, float *slope) {
  for (int32_t c = 0; c < n; c++) {
    if (outflow[c] != 0 && (ldd[c] != 5 && ldd[c] != -1))
      slope[c] = (alt[c] - alt[outflow[c] - 1]) /
                 (1000.0 * dist[(ldd[c] - 1) * 360 + row[c] - 1] * ((meander[c] + meander[outflow[c] - 1]) / 2));
    else
      slope[c] = 0;
  }
  {
    float minslope = 0.00001;
    for (int32_t c = 0; c < n; c++)
      if (slope[c] < minslope) slope[c] = minslope;
  }
}

/* rout_prepare.cpp:790-832 */
void wgo_river_length(int32_t n, const float *meander, const int16_t *row, const int16_t *col, const float *dist,
                      const int32_t *outflow, float *len) {
  for (int32_t c = 0; c < n; c++) {
    int32_t o = outflow[c] - 1;
    if (o == -1)
      len[c] = 55.;
    else if (row[c] == row[o])
      len[c] = dist[3 * 360 + row[c] - 1];
    else {
      if (row[c] > row[o]) {
        if (col[c] == col[o])
          len[c] = dist[7 * 360 + row[c] - 1];
        else
          len[c] = dist[8 * 360 + row[c] - 1];
      } else {
        if (col[c] == col[o])
          len[c] = dist[1 * 360 + row[c] - 1];
        else
          len[c] = dist[2 * 360 + row[c] - 1];
      }
    }
    if (o == -1)
      len[c] *= meander[c];
    else
      len[c] *= ((meander[c] + meander[o]) / 2);
  }
}

/* rout_prepare.cpp:895-944 */
void wgo_alloc_coeff(int32_t n, const int32_t *outflow, const double *res_area, const double *mean_outflow,
                     double *alloc) {
  double *inflow_res = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  for (int32_t c = 0; c < n; c++)
    for (int i = 0; i < 5; i++) alloc[(int64_t)c * 5 + i] = 1.0;
  for (int32_t c = 0; c < n; c++) {
    if (res_area[c] > 0) {
      short i = 0;
      int32_t d = outflow[c];
      while (i < 5 && d > 0 && res_area[d - 1] <= 0) {
        inflow_res[d - 1] += mean_outflow[c];
        d = outflow[d - 1];
        i++;
      }
    }
  }
  for (int32_t c = 0; c < n; c++) {
    if (res_area[c] > 0) {
      short i = 0;
      int32_t d = outflow[c];
      while (i < 5 && d > 0 && res_area[d - 1] <= 0) {
        alloc[(int64_t)c * 5 + i] = mean_outflow[c] / inflow_res[d - 1];
        d = outflow[d - 1];
        i++;
      }
    }
  }
  free(inflow_res);
}

/* rout_prepare.cpp:950-1019 (variante 3); a never-assigned start is taken as 0 */
void wgo_reservoir_start_month(int32_t n, const double *monthly, const double *annual, int8_t *start_month) {
  for (int32_t c = 0; c < n; c++) {
    const double *q = monthly + (int64_t)c * 12;
    int start = 0, dry_length = 0, start_ = 0, counter_length = 0, last_dry = 1;
    int beg_month = 0, month = 0;
    while (month < 12 && q[month] < annual[c]) month++;
    beg_month = month;
    for (int m = 0; m < 12; m++) {
      month = m + beg_month;
      if (month > 11) month -= 12;
      if (q[month] < annual[c]) {
        counter_length++;
        if (!last_dry) {
          start_ = month;
          last_dry = 1;
        }
      } else if (last_dry) {
        last_dry = 0;
        if (counter_length > dry_length) {
          start = start_;
          dry_length = counter_length;
          counter_length = 0;
        }
      }
    }
    if (counter_length > dry_length) {
      start = start_;
      dry_length = counter_length;
    }
    start_month[c] = (int8_t)(start + 1);
  }
}
