/*
 * wg_oracle.c -- CPU oracle (test infrastructure only, see wg_oracle.h).
 *
 * Plain C99 restatement of the reference's arithmetic for the parity option
 * profile (time_series=1, cloud=1, intercept=1, calc_albedo=0, petOpt=0,
 * use_kc=0, riverveloOpt=2, subtract_use=0, resOpt=1 with resYearOpt=1,
 * statcorrOpt=0, aridareaOpt=1, fractionalRoutingOpt=1, riverEvapoOpt=1,
 * clclOpt=0, calc_wtemp=0, glacierOpt=0, grid_store=0, timeStepsPerDay=1).
 * Expressions keep the reference's operand order and association so that IEEE
 * rounding matches; build with -ffp-contract=off and without fast-math.
 * Reference paths are relative to source/ of HydrologyFrankfurt/WaterGAP2.
 */
#include "wg_oracle.h"

#include <math.h>
/* The C library functions the reference calls.  By default the deterministic
 * exp/pow of wg_libm.h (correctly rounded in all sampled cases, the same bits on
 * the CPU and the GPU); built with -DWG_ORACLE_GLIBC the host C library's. */
#include "../watergap2_amd/csrc/wg_libm.h"
#ifdef WG_ORACLE_GLIBC
#define EXP exp
#define POW pow
#else
#define EXP wg_exp
#define POW wg_pow
#endif
#include <stdlib.h>
#include <string.h>

/* daily.cpp:1866-1880 setParameters */
#define CANOPY_EVAPO_EXP 0.66666666
#define RUNOFF_FRAC_BUILTUP 0.5
#define OPEN_WATER_ALBEDO 0.08
/* routing.cpp:124-129 setParameters, routing.cpp:1662 Kswbgw, routing.h:24 minStorVol */
#define LAKE_OUTFLOW_EXP 1.5
#define WETL_OUTFLOW_EXP 2.5
#define EVAPO_RED_EXP 3.32193
#define EVAPO_RED_EXP_RES 2.81383
#define KSWBGW 10.
#define MIN_STOR_VOL 1.e-15

/* routing.cpp:1636 (julian day of the first day of each month) */
static const short first_day_in_month_r[12] = {1, 32, 60, 91, 121, 152, 182, 213, 244, 274, 305, 335};
/* integrateWGHM.cpp:100 */
static const short days_in_month[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};

struct wgo_model {
  int32_t n;
  int threads;             /* OpenMP threads of the vertical loop */
  /* statics (reference types) */
  const wg_static_fields *st;
  const wg_lct_tables *lct;
  double *gamma, *cfa;     /* G_gammaHBV, G_cellCorrFact (integrateWGHM.cpp:318-366) */
  float *smax;             /* G_Smax */
  float *lai_max;          /* lai.cpp:78-81 */
  float lai_a[WG_NLCT], lai_b[WG_NLCT]; /* lai.cpp:86-89 */
  short *contcell;         /* geo.cpp:54-62 */
  double *river_length, *river_slope, *roughness, *width_bf, *bottom_width, *river_stor_max;
  double *lake_depth, *wetl_depth;   /* routing.cpp:5621-5644 */
  double *f_glo_lake, *fswb_init;    /* routing.cpp:745-765 */
  double *mean_demand;               /* routing.cpp:360-403, m3/s */
  int32_t *routing_cell;              /* routing.cpp:530-538 */
  /* vertical state (daily.h) */
  double *canopy, *soil, *snow, *snow_elev, *transfer;
  int32_t *lai_days, *lai_status;
  double *lai_precsum;
  /* hand-off (daily.h:56-78) */
  double *surf_runoff, *gw_recharge, *ow_prec, *ow_pet, *aet_uncorr, *prec_out;
  /* routing state */
  double *gw, *loc_lake, *loc_wetl, *glo_lake, *glo_wetl, *res, *river;
  double *red_ll, *red_lw, *red_gl, *red_gw, *red_res, *red_river;
  double *river_frac_next, *river_frac_change, *fswb_next, *laf, *laf_prev, *laf_next, *k_release;
  double *river_inflow, *discharge;
  double *pot_acc;         /* G_potCellRunoff (routing.h:286), km3 since the last annual init */
  /* water use (subtract_use 2/3, use_alloc 0, aggrNUsGloLakResOpt 0) */
  int use_on, use_delayed;
  double *use_nus, *use_nug, *use_wu, *use_cu; /* [N][12] daily values of the year (dailyNUInit) */
  double *use_frgi;        /* G_fractreturngw_irrig */
  double *use;             /* [WG_USE_NFIELDS][N] state */
  int32_t started;
  /* stations */
  int32_t n_st;
  int32_t *st_cell;
  int32_t *st_slot;        /* first station of each cell, -1 = none (index map) */
  int station_scan;        /* 1: the reference's scan of all stations per cell */
  int aggr_scan;           /* 1: the O(N) riparian scan of aggrNUsGloLakResOpt = 1 */
  int64_t aggr_matches;
  double *st_sum;
};

int wgo_month_of_day(int day) {
  int m = 0;
  while (m < 11 && day > first_day_in_month_r[m + 1] - 1) m++;
  return m;
}

/* ------------------------------------------------------------------ */
/* routing schedule: rout_prepare.cpp                                   */
/* ------------------------------------------------------------------ */

/* rout_prepare.cpp:338-439 inflow_cells and 619-670 outflow_cell.  cell_at is the
 * GCRC lattice (row-major 360 x 720, 1-based cell, 0 = no land); GCRC_2 adds a
 * zero row above row 1 and below row 360 and wraps longitude (rout_prepare.cpp:132-146). */
static int32_t gcrc2(const int32_t *cell_at, int col, int row) {
  if (row < 1 || row > 360) return 0;
  if (col < 1) col += 720;
  if (col > 720) col -= 720;
  return cell_at[(row - 1) * 720 + (col - 1)];
}

int wgo_inflow_outflow(int32_t n, const int8_t *ldd, const int16_t *row, const int16_t *col,
                       const int32_t *cell_at, int32_t *inflow9, int32_t *outflow) {
  for (int32_t i = 0; i < n; i++) {
    int r = row[i], c = col[i];
    int32_t *in = inflow9 + (int64_t)i * 9;
    int32_t g;
    for (int k = 0; k < 9; k++) in[k] = 0;
    g = gcrc2(cell_at, c + 1, r - 1); if (g && ldd[g - 1] == 1) in[8] = g;
    g = gcrc2(cell_at, c, r - 1);     if (g && ldd[g - 1] == 2) in[7] = g;
    g = gcrc2(cell_at, c - 1, r - 1); if (g && ldd[g - 1] == 3) in[6] = g;
    g = gcrc2(cell_at, c + 1, r);     if (g && ldd[g - 1] == 4) in[5] = g;
    g = gcrc2(cell_at, c - 1, r);     if (g && ldd[g - 1] == 6) in[3] = g;
    g = gcrc2(cell_at, c + 1, r + 1); if (g && ldd[g - 1] == 7) in[2] = g;
    g = gcrc2(cell_at, c, r + 1);     if (g && ldd[g - 1] == 8) in[1] = g;
    g = gcrc2(cell_at, c - 1, r + 1); if (g && ldd[g - 1] == 9) in[0] = g;
    switch (ldd[i]) {
      case 1: outflow[i] = gcrc2(cell_at, c - 1, r + 1); break;
      case 2: outflow[i] = gcrc2(cell_at, c, r + 1); break;
      case 3: outflow[i] = gcrc2(cell_at, c + 1, r + 1); break;
      case 4: outflow[i] = gcrc2(cell_at, c - 1, r); break;
      case 6: outflow[i] = gcrc2(cell_at, c + 1, r); break;
      case 7: outflow[i] = gcrc2(cell_at, c - 1, r - 1); break;
      case 8: outflow[i] = gcrc2(cell_at, c, r - 1); break;
      case 9: outflow[i] = gcrc2(cell_at, c + 1, r - 1); break;
      default: outflow[i] = 0; break;
    }
  }
  return 0;
}

/* rout_prepare.cpp:834-886: repeated sweeps; each sweep numbers every cell whose
 * remaining inflow count is zero, in ascending cell index, then decrements the
 * count of the cell itself and of its outflow cell.  A cell never numbered means a
 * cycle or an inconsistent graph (rout_prepare.cpp:876-881). */
int wgo_rout_order(int32_t n, const int32_t *inflow9, const int32_t *outflow, int32_t *order,
                   int32_t *level) {
  int32_t *routing = (int32_t *)calloc((size_t)n, sizeof(int32_t));
  char *list = (char *)calloc((size_t)n, 1);
  int32_t next = 1, sweep = 0, counter_step;
  for (int32_t i = 0; i < n; i++) {
    order[i] = -99;
    for (int k = 0; k < 9; k++)
      if (inflow9[(int64_t)i * 9 + k] > 0) routing[i]++;
  }
  do {
    counter_step = 0;
    for (int32_t i = 0; i < n; i++) {
      if (routing[i] == 0) {
        order[i] = next++;
        if (level) level[i] = sweep;
        counter_step++;
        list[i] = 1;
      } else {
        list[i] = 0;
      }
    }
    for (int32_t i = 0; i < n; i++) {
      if (list[i]) {
        routing[i]--;
        if (outflow[i] > 0) routing[outflow[i] - 1]--;
      }
    }
    sweep++;
  } while (counter_step > 0);
  int bad = 0;
  for (int32_t i = 0; i < n; i++)
    if (routing[i] >= 0 || order[i] == -99) bad = 1;
  free(routing);
  free(list);
  return bad ? WG_E_CYCLE : 0;
}

/* ------------------------------------------------------------------ */
/* initialisation                                                       */
/* ------------------------------------------------------------------ */

#define ALLOC_D(p) (p) = (double *)calloc((size_t)n, sizeof(double))

/* routing.cpp:360-403: mean demand of the cell and of up to reservoir_dsc = 5
 * downstream cells (m3/yr -> m3/s), from the reservoirs in operation. */
static void mean_demand(wgo_model *m, const wg_static_fields *st) {
  const int32_t n = m->n;
  for (int32_t i = 0; i < n; i++) {
    m->mean_demand[i] = 0;
    if (st->mean_nus) {
      short k = 0;
      int32_t downstreamCell = st->downstream[i];
      m->mean_demand[i] += st->mean_nus[i];
      while (k < 5 && downstreamCell > 0 &&
             (downstreamCell < n ? st->reservoir_area[downstreamCell] : 0.) <= 0) {
        const double coeff = st->alloc_coeff ? st->alloc_coeff[(size_t)i * 5 + k] : 1.0;
        k++;
        m->mean_demand[i] += st->mean_nus[downstreamCell - 1] * coeff;
        downstreamCell = st->downstream[downstreamCell - 1];
      }
    }
    m->mean_demand[i] = m->mean_demand[i] / 31536000.;
  }
}

wgo_model *wgo_create(int32_t n, const wg_static_fields *st) {
  wgo_model *m = (wgo_model *)calloc(1, sizeof(wgo_model));
  if (m) m->threads = 1;
  const double *par = st->params;
  m->n = n;
  m->st = st;
  m->lct = st->lct;
  ALLOC_D(m->gamma); ALLOC_D(m->cfa);
  m->smax = (float *)calloc((size_t)n, sizeof(float));
  m->lai_max = (float *)calloc((size_t)n, sizeof(float));
  m->contcell = (short *)calloc((size_t)n, sizeof(short));
  ALLOC_D(m->river_length); ALLOC_D(m->river_slope); ALLOC_D(m->roughness); ALLOC_D(m->width_bf);
  ALLOC_D(m->bottom_width); ALLOC_D(m->river_stor_max); ALLOC_D(m->lake_depth); ALLOC_D(m->wetl_depth);
  ALLOC_D(m->f_glo_lake); ALLOC_D(m->fswb_init); ALLOC_D(m->mean_demand);
  m->routing_cell = (int32_t *)calloc((size_t)n, sizeof(int32_t));
  ALLOC_D(m->canopy); ALLOC_D(m->soil); ALLOC_D(m->snow); ALLOC_D(m->transfer);
  m->snow_elev = (double *)calloc((size_t)n * WG_NELEV, sizeof(double));
  m->lai_days = (int32_t *)calloc((size_t)n, sizeof(int32_t));
  m->lai_status = (int32_t *)calloc((size_t)n, sizeof(int32_t));
  ALLOC_D(m->lai_precsum);
  ALLOC_D(m->surf_runoff); ALLOC_D(m->gw_recharge); ALLOC_D(m->ow_prec); ALLOC_D(m->ow_pet);
  ALLOC_D(m->aet_uncorr); ALLOC_D(m->prec_out);
  ALLOC_D(m->gw); ALLOC_D(m->loc_lake); ALLOC_D(m->loc_wetl); ALLOC_D(m->glo_lake); ALLOC_D(m->glo_wetl);
  ALLOC_D(m->res); ALLOC_D(m->river);
  ALLOC_D(m->red_ll); ALLOC_D(m->red_lw); ALLOC_D(m->red_gl); ALLOC_D(m->red_gw); ALLOC_D(m->red_res);
  ALLOC_D(m->red_river);
  ALLOC_D(m->river_frac_next); ALLOC_D(m->river_frac_change); ALLOC_D(m->fswb_next); ALLOC_D(m->laf);
  ALLOC_D(m->laf_prev); ALLOC_D(m->laf_next); ALLOC_D(m->k_release);
  ALLOC_D(m->river_inflow); ALLOC_D(m->discharge); ALLOC_D(m->pot_acc);

  /* lai.cpp:86-89: lai_factor_a = 0.1*decid, lai_factor_b = (1-decid)*evergreen (float) */
  for (int i = 0; i < WG_NLCT; i++) {
    m->lai_a[i] = (float)(0.1 * m->lct->decid_frac[i]);
    m->lai_b[i] = (1 - m->lct->decid_frac[i]) * m->lct->evergreen_red[i];
  }
  for (int32_t i = 0; i < n; i++) {
    /* integrateWGHM.cpp:318-366 gamma and CFA from the parameter set */
    m->gamma[i] = par[(size_t)WG_P_GAMRUN_C * n + i];
    m->cfa[i] = par[(size_t)WG_P_CFA * n + i];
    m->smax[i] = st->smax[i];
    /* lai.cpp:78-81 */
    m->lai_max[i] = (float)(par[(size_t)WG_M_LAI * n + i] * m->lct->lai[st->land_cover[i] - 1]);
    /* geo.cpp:54-62 (limit_contcell_pct = 1e-6, geo.h) */
    m->contcell[i] = (st->contfreq[i] < 0.000001) ? 0 : 1;
    /* routing.cpp:476-507 */
    m->river_length[i] = (double)st->river_length[i];
    m->river_length[i] *= st->contfreq[i] / 100.;
    m->river_slope[i] = (double)st->river_slope[i];
    m->roughness[i] = (double)st->roughness[i];
    {
      double bf = (double)st->bankfull_flow[i];
      double depth_bf;
      if (bf < 0.05) bf = 0.05;
      m->width_bf[i] = 2.71 * POW(bf, 0.557);
      depth_bf = 0.349 * POW(bf, 0.341);
      m->bottom_width[i] = m->width_bf[i] - 2.0 * 2.0 * depth_bf;
      m->river_stor_max[i] = m->river_length[i] * 0.5 * depth_bf / 1000. *
                             (m->bottom_width[i] / 1000. + m->width_bf[i] / 1000.);
    }
    /* routing.cpp:5621-5644 (CONV_FACTOR__LENGTH__M_TO_KM = 0.001, conversion.h:6) */
    m->lake_depth[i] = (par[(size_t)WG_P_LAK_D * n + i] * 0.001);
    m->wetl_depth[i] = (par[(size_t)WG_P_WET_D * n + i] * 0.001);
    /* routing.cpp:745-765 initFractionStatus */
    {
      double fll = st->loc_lake[i] / 100., flw = st->loc_wetland[i] / 100.;
      double fgw = st->glo_wetland[i] / 100.;
      m->f_glo_lake[i] = st->glo_lake[i] / 100.;
      m->fswb_init[i] = fll + flw + fgw;
      m->fswb_next[i] = m->fswb_init[i];
    }
    /* routing.cpp:789-818 setStoragesToZero with riverveloOpt=2 */
    m->river[i] = m->river_stor_max[i];
    m->gw[i] = 0.;
    /* routing.cpp:5647-5720 setLakeWetlToMaximum, resYearOpt=1, all reservoirs in operation */
    m->loc_lake[i] = (st->loc_lake[i] / 100.) * st->cell_area[i] * m->lake_depth[i];
    m->loc_wetl[i] = (st->loc_wetland[i] / 100.) * st->cell_area[i] * m->wetl_depth[i];
    m->glo_lake[i] = st->lake_area[i] * m->lake_depth[i];
    m->glo_wetl[i] = (st->glo_wetland[i] / 100.) * st->cell_area[i] * m->wetl_depth[i];
    m->red_res[i] = 1.;
    m->res[i] = st->stor_cap[i];
    m->red_ll[i] = 1.;
    m->red_lw[i] = 1.;
    m->red_gl[i] = 1.;
    m->red_gw[i] = 1.;
    m->red_river[i] = 0.5;
    m->river_frac_next[i] = m->red_river[i] * m->river_length[i] * m->width_bf[i] / 1000. / st->cell_area[i];
    /* routing.cpp:405-406 */
    m->k_release[i] = 0.1;
    /* routing.cpp:1181-1205 annualInit, first year */
    m->laf[i] = (st->contfreq[i] - (st->glo_lake[i] + st->glo_wetland[i] + st->loc_lake[i] +
                                   st->loc_wetland[i] + st->glo_res[i]));
    if (m->laf[i] < 0.) m->laf[i] = 0.;
    m->laf_prev[i] = 0.;
  }
  /* routing.cpp:360-403: mean demand of the cell and of up to reservoir_dsc = 5
   * downstream cells (m3/yr -> m3/s).  G_downstreamCell holds 1-based cell numbers
   * and the loop tests G_reservoir_area_full[downstreamCell] with that number,
   * i.e. the cell after the downstream cell; an index past the last cell is taken
   * as a cell without a reservoir.  Allocation coefficients default to 1
   * (rout_prepare.cpp:906-909). */
  mean_demand(m, st);
  /* routing.cpp:530-538 */
  for (int32_t i = 0; i < n; i++) m->routing_cell[st->rout_order[i] - 1] = i + 1;
  m->started = 0;
  return m;
}

/* routingClass::annualInit for a later year (routing.cpp:979-1495), parity profile:
 * st holds this year's reservoir grids (Parts 1-2, routing.cpp:1015-1100, as the
 * caller derives them); Part 4 (routing.cpp:1300-1421) moves the land of a growing
 * reservoir fraction into the reservoir storage of the cell's outflow cell
 * assignment.  The branch of routing.cpp:1354-1357 (land not shrunk against the
 * day before) is refused with -5, as the library refuses it. */
int wgo_annual_init(wgo_model *m, const wg_static_fields *st, const int32_t *outflow_cell, int apply_changes,
                    double *glo_res_prev) {
  const int32_t n = m->n;
  if (apply_changes) {
    for (int32_t i = 0; i < n; i++) {
      double change = st->glo_res[i] - glo_res_prev[i];           /* routing.cpp:1315 */
      if (change > 0.) {
        double laf = (glo_res_prev[i] > 0.) ? m->laf[i] - change : m->laf[i] - st->glo_res[i];
        if (laf < 0.) laf = 0.;
        if (laf - m->laf_prev[i] >= 0.) return -5;
      }
    }
    for (int32_t i = 0; i < n; i++) {
      double change = st->glo_res[i] - glo_res_prev[i];
      if (change > 0.) {
        /* routing.cpp:1326-1332 */
        if (glo_res_prev[i] > 0.)
          m->laf[i] = m->laf[i] - change;
        else
          m->laf[i] = m->laf[i] - st->glo_res[i];
        if (m->laf[i] < 0.) m->laf[i] = 0.;
        {
          /* routing.cpp:1346-1400 */
          double laf_change = m->laf[i] - m->laf_prev[i];
          double cellArea = st->cell_area[i];
          double canopy_km3 = m->canopy[i] * cellArea / 1000000.0 * ((-laf_change) / 100.0);
          double soil_km3 = m->soil[i] * cellArea / 1000000.0 * ((-laf_change) / 100.0);
          double snow_km3 = m->snow[i] * cellArea / 1000000.0 * ((-laf_change) / 100.0);
          m->res[outflow_cell[i] - 1] += canopy_km3 + soil_km3 + snow_km3;
        }
      }
    }
  }
  /* this year's reservoirs: the statics, and the demand terms that depend on them
   * (routing.cpp:360-403) */
  m->st = st;
  mean_demand(m, st);
  for (int32_t i = 0; i < n; i++) m->pot_acc[i] = 0.;                /* routing.cpp:1009 */
  for (int32_t i = 0; i < n; i++) glo_res_prev[i] = st->glo_res[i];  /* routing.cpp:5479-5483 at year end */
  return 0;
}

void wgo_destroy(wgo_model *m) {
  if (!m) return;
  double **d[] = {&m->gamma, &m->cfa, &m->river_length, &m->river_slope, &m->roughness, &m->width_bf,
                  &m->bottom_width, &m->river_stor_max, &m->lake_depth, &m->wetl_depth, &m->f_glo_lake,
                  &m->fswb_init, &m->mean_demand, &m->canopy, &m->soil, &m->snow, &m->snow_elev, &m->transfer,
                  &m->lai_precsum, &m->surf_runoff, &m->gw_recharge, &m->ow_prec, &m->ow_pet,
                  &m->aet_uncorr, &m->prec_out, &m->gw, &m->loc_lake, &m->loc_wetl, &m->glo_lake,
                  &m->glo_wetl, &m->res, &m->river, &m->red_ll, &m->red_lw, &m->red_gl, &m->red_gw,
                  &m->red_res, &m->red_river, &m->river_frac_next, &m->river_frac_change,
                  &m->fswb_next, &m->laf, &m->laf_prev, &m->laf_next, &m->k_release,
                  &m->river_inflow, &m->discharge, &m->st_sum, &m->pot_acc, &m->use_nus, &m->use_nug,
                  &m->use_wu, &m->use_cu, &m->use_frgi, &m->use};
  for (size_t i = 0; i < sizeof(d) / sizeof(d[0]); i++) free(*d[i]);
  free(m->smax); free(m->lai_max); free(m->contcell); free(m->routing_cell);
  free(m->lai_days); free(m->lai_status); free(m->st_cell); free(m->st_slot);
  free(m);
}

void wgo_set_params(wgo_model *m, const double *gamma, const double *cfa, const float *smax) {
  size_t nb = (size_t)m->n;
  if (gamma) memcpy(m->gamma, gamma, nb * sizeof(double));
  if (cfa) memcpy(m->cfa, cfa, nb * sizeof(double));
  if (smax) memcpy(m->smax, smax, nb * sizeof(float));
}

/* ------------------------------------------------------------------ */
/* LAI: lai.cpp:152-306                                                 */
/* ------------------------------------------------------------------ */

/* lai.cpp:179-241 dailyLai_growing */
static double lai_growing(int *days, short init_days, int *status, char lct, int arid, double lmin,
                          double lmax, double *precsum, double prec) {
  if (*status == 0) {
    if (*days >= init_days) {
      (*days)++;
      (*precsum) += prec;
      if (*precsum > 40.) {
        if (*days >= init_days + 30) {
          *days = init_days + 30;
          *status = 1;
        }
        return (lmin + (lmax - lmin) * (*days - init_days) / 30.);
      } else {
        *days = init_days;
        return (lmin);
      }
    } else {
      (*days)++;
      (*precsum) += prec;
      return (lmin);
    }
  } else {
    if (*days <= 30) {
      (*days)--;
      if (lct <= 2) *status = 0;
      if (*days <= 0) {
        *days = 0;
        *status = 0;
        *precsum = 0.;
      }
      return (lmax - (lmax - lmin) * (30 - *days) / 30.);
    } else {
      if ((arid) && (prec < 0.5))
        (*days)--;
      else
        *days = 30 + init_days;
      return (lmax);
    }
  }
}

/* lai.cpp:243-293 dailyLai_nogrowing */
static double lai_nogrowing(int *days, short init_days, int *status, double lmin, double lmax,
                            double *precsum, double prec) {
  if (*status == 0) {
    if (*days > init_days) {
      (*days)++;
      (*precsum) += prec;
      if (*precsum > 40.) {
        if (*days >= init_days + 30) {
          *days = init_days + 30;
          *status = 1;
        }
        return (lmin + (lmax - lmin) * (*days - init_days) / 30.);
      } else {
        *days = init_days;
        return (lmin);
      }
    } else {
      (*precsum) += prec;
      return (lmin);
    }
  } else {
    if (*days <= 30) {
      (*days)--;
      if (*days <= 0) {
        *days = 0;
        *status = 0;
        *precsum = 0.;
      }
      return (lmax - (lmax - lmin) * (30 - *days) / 30.);
    } else {
      (*days)--;
      return (lmax);
    }
  }
}

/* ------------------------------------------------------------------ */
/* vertical balance: daily.cpp:94-1279 for one cell                     */
/* ------------------------------------------------------------------ */

static void vertical_cell(wgo_model *m, int32_t i, double tempC_f, double prec_f, double sw_f,
                          double lw_f) {
  const wg_static_fields *st = m->st;
  const int32_t n = m->n;
  const double *par = st->params;
#define P(id) (par[(size_t)(id) * n + i])
  const int lct = st->land_cover[i];
  int arid_gw = 0;
  double dailyTempC, dailyPrec, dailyShortWave, dailyLongWave;
  double alpha, maxDailyPET, albedo, lat_heat;
  double landAreaFrac, lafPrev;

  /* daily.cpp:159-169: land area fraction of today and of the previous step */
  landAreaFrac = m->laf[i];
  lafPrev = m->started ? m->laf_prev[i] : landAreaFrac;

  /* daily.cpp:231-248 */
  dailyPrec = prec_f;
  dailyTempC = tempC_f;
  dailyShortWave = sw_f;
  dailyLongWave = lw_f;
  dailyPrec = P(WG_M_PREC) * dailyPrec;
  double temp2 = dailyTempC + 237.3;
  double e_s = 0.6108 * EXP(17.27 * dailyTempC / temp2);

  /* daily.cpp:329-349 arid/humid from G_aindex */
  if (st->aindex[i] == 1) {
    alpha = P(WG_P_PTC_ARI);
    maxDailyPET = P(WG_P_PET_MXDY);
    arid_gw = 1;
  } else {
    alpha = P(WG_P_PTC_HUM);
    maxDailyPET = P(WG_P_PET_MXDY);
  }

  /* daily.cpp:355 -> lai.cpp:152-177 getDailyLai */
  double dailyLai;
  {
    double lmax = (double)m->lai_max[i];
    double lmin = m->lai_a[lct - 1] + m->lai_b[lct - 1] * m->lai_max[i];
    if (dailyTempC > 8.)
      dailyLai = lai_growing(&m->lai_days[i], m->lct->initial_days[lct - 1], &m->lai_status[i],
                             (char)lct, arid_gw, lmin, lmax, &m->lai_precsum[i], dailyPrec);
    else
      dailyLai = lai_nogrowing(&m->lai_days[i], m->lct->initial_days[lct - 1], &m->lai_status[i],
                               lmin, lmax, &m->lai_precsum[i], dailyPrec);
  }

  /* daily.cpp:366-378 albedo (calc_albedo=0, use_kc=0) */
  if (m->snow[i] > 3.)
    albedo = m->lct->albedo_snow[lct - 1];
  else
    albedo = m->lct->albedo[lct - 1];

  /* daily.cpp:381-386 */
  if (dailyTempC > 0)
    lat_heat = 2.501 - 0.002361 * dailyTempC;
  else
    lat_heat = 2.835;

  /* daily.cpp:404-462 (cloud=1); calc_ext_rad only feeds cloud=2 and outputs */
  double conv_Wm2_to_mmd = 0.0864 / lat_heat;
  double solar_rad = conv_Wm2_to_mmd * dailyShortWave;
  double long_wave_rad_in = conv_Wm2_to_mmd * dailyLongWave;
  double emissivity = m->lct->emissivity[lct - 1];
  double temp_K = dailyTempC + 273.2;
  const double stefan_boltz_const = 0.000000004903;
  double long_wave_rad_out = emissivity * stefan_boltz_const * POW(temp_K, 4.) / lat_heat;
  double net_long_wave_rad = long_wave_rad_in - long_wave_rad_out;
  double net_short_wave_rad = solar_rad * (1. - albedo);
  double net_rad = P(WG_M_NETRAD) * (net_short_wave_rad + net_long_wave_rad);
  double openWaterNetShortWaveRad = solar_rad * (1. - OPEN_WATER_ALBEDO);
  double openWaterNetRad = openWaterNetShortWaveRad + net_long_wave_rad;

  /* daily.cpp:476-522 Priestley-Taylor */
  double dailyPET, dailyOpenWaterPET;
  double inc_svp = 4098. * e_s / (temp2 * temp2);
  double c3 = 0.0016286 * 101.3;
  double gamma = c3 / lat_heat;
  if (net_rad <= 0.)
    dailyPET = 0.;
  else
    dailyPET = alpha * (inc_svp * net_rad) / (inc_svp + gamma);
  if (openWaterNetRad <= 0.)
    dailyOpenWaterPET = 0.;
  else
    dailyOpenWaterPET = alpha * (inc_svp * openWaterNetRad) / (inc_svp + gamma);

  /* daily.cpp:773-777 */
  m->ow_prec[i] = dailyPrec;
  m->ow_pet[i] = dailyOpenWaterPET;
  m->prec_out[i] = dailyPrec;

  double landStorageChangeSum = 0., initialStorage = 0.;
  double max_canopy_storage = 0., dailyCanopyEvapo = 0., daily_prec_to_soil = 0.;
  double dailySoilPET = 0., canopy_water_content = 0.;
  double snowmelt = 0., dailySnowEvapo = 0., dailyEffPrec = 0.;
  double immediate_runoff = 0., dailyAET = 0., daily_runoff = 0., total_daily_runoff = 0.;
  double daily_gw_recharge = 0., pot_gw_recharge = 0.;
  double soil_saturation = 0., soil_water_overflow = 0., neg_land_aet = 0., neg_runoff = 0.;
  double dailyCellLandAET = 0., dailyCellLandAET_uncorr = 0.;

  const double T_SNOWFZ = P(WG_P_T_SNOWFZ), T_SNOWMT = P(WG_P_T_SNOWMT);
  const double T_GRADNT = P(WG_P_T_GRADNT), DEGDAY_F = P(WG_M_DEGDAY_F);

  /* daily.cpp:825-895 interception */
  if (landAreaFrac <= 0.) {
    m->transfer[i] = m->canopy[i];
    m->canopy[i] = 0.;
    dailyCanopyEvapo = 0.;
  } else {
    m->canopy[i] *= lafPrev / landAreaFrac;
    if (fabs(m->canopy[i]) <= MIN_STOR_VOL) m->canopy[i] = 0.;
    initialStorage = m->canopy[i];
    if (dailyLai > 0.00001) {
      double canopy_deficiency;
      max_canopy_storage = P(WG_P_MCWH) * dailyLai;
      canopy_deficiency = max_canopy_storage - m->canopy[i];
      if (dailyPrec < canopy_deficiency) {
        m->canopy[i] += dailyPrec;
        daily_prec_to_soil = 0.;
      } else {
        m->canopy[i] = max_canopy_storage;
        daily_prec_to_soil = dailyPrec - canopy_deficiency;
      }
      canopy_water_content = m->canopy[i];
      dailyCanopyEvapo = dailyPET * POW((canopy_water_content / max_canopy_storage), CANOPY_EVAPO_EXP);
      if (dailyCanopyEvapo > canopy_water_content) {
        dailyCanopyEvapo = canopy_water_content;
        dailySoilPET = dailyPET - canopy_water_content;
        m->canopy[i] = 0.0;
      } else {
        m->canopy[i] -= dailyCanopyEvapo;
        dailySoilPET = dailyPET - dailyCanopyEvapo;
      }
    } else {
      daily_prec_to_soil = dailyPrec;
      dailySoilPET = dailyPET;
      dailyCanopyEvapo = 0.0;
    }
    landStorageChangeSum += m->canopy[i] - initialStorage;
  }
  if (dailySoilPET < 0.) dailySoilPET = 0.0;

  /* daily.cpp:897-1062 snow on 100 elevation sub-cells */
  double daily_snow_to_soil = 0., TempElevMax = 0., snowStorageChange = 0.;
  int thresh_elev = 0;
  const int16_t *elev = st->elevation + (size_t)i * WG_NELEV;
  double *se = m->snow_elev + (size_t)i * WG_NELEV;
  m->snow[i] = 0.;
  for (int e = 1; e < 101; e++) {
    double temp_elev, daily_snow_to_soil_elev = 0., dailyEffPrec_elev = 0., snowmelt_elev = 0.;
    double dailyEffPrecBeforeSnowMelt_elev = 0.;
    temp_elev = dailyTempC - ((elev[e] - elev[0]) * T_GRADNT);
    if (landAreaFrac <= 0.) {
      m->transfer[i] += se[e] / 100.;
      se[e] = 0.;
      m->snow[i] = 0.;
    } else {
      se[e] = se[e] * lafPrev / landAreaFrac;
      if (fabs(se[e]) <= MIN_STOR_VOL) se[e] = 0.;
      initialStorage = se[e];
      if (se[e] > 1000.) {
        if (thresh_elev == 0.) {
          thresh_elev = elev[e];
        } else if (thresh_elev > 0.) {
          temp_elev = dailyTempC - ((thresh_elev - elev[0]) * T_GRADNT);
        }
      }
      if (temp_elev <= T_SNOWFZ) {
        dailyEffPrecBeforeSnowMelt_elev = 0.;
        daily_snow_to_soil_elev = daily_prec_to_soil;
        se[e] += daily_snow_to_soil_elev;
        if (se[e] > dailySoilPET) {
          se[e] -= dailySoilPET;
          dailySnowEvapo += dailySoilPET;
        } else {
          dailySnowEvapo += se[e];
          se[e] = 0.;
        }
      } else {
        dailyEffPrecBeforeSnowMelt_elev = daily_prec_to_soil;
      }
      if (temp_elev > T_SNOWMT) {
        if (!(se[e] < 0.)) {
          snowmelt_elev = DEGDAY_F * m->lct->ddf[lct - 1] * (temp_elev - T_SNOWMT);
          if (snowmelt_elev > se[e]) {
            snowmelt_elev = se[e];
            se[e] = 0.;
          } else {
            se[e] -= snowmelt_elev;
          }
        }
      }
      dailyEffPrec_elev = dailyEffPrecBeforeSnowMelt_elev + snowmelt_elev;
      snowStorageChange += se[e] - initialStorage;
      if (e == 1) TempElevMax = temp_elev;
      m->snow[i] += se[e];
      dailyEffPrec += dailyEffPrec_elev;
      snowmelt += snowmelt_elev;
      daily_snow_to_soil += daily_snow_to_soil_elev;
    }
  }
  if (landAreaFrac > 0.) {
    m->snow[i] /= 100.;
    dailyEffPrec /= 100.;
    daily_snow_to_soil /= 100.;
    snowmelt /= 100.;
    dailySnowEvapo /= 100.;
    snowStorageChange /= 100.;
    landStorageChangeSum += snowStorageChange;
  }

  /* daily.cpp:1066-1071 immediate runoff over built-up area */
  if (st->built_up[i] > 0.) {
    immediate_runoff = RUNOFF_FRAC_BUILTUP * dailyEffPrec * st->built_up[i];
    dailyEffPrec -= immediate_runoff;
  }

  /* daily.cpp:1080-1239 soil, runoff, AET, recharge */
  const double cfa = m->cfa[i];
  const float smax = m->smax[i];
  if (landAreaFrac <= 0.) {
    m->transfer[i] += m->soil[i];
    m->transfer[i] *= cfa;
    m->soil[i] = 0.;
    daily_gw_recharge = 0.;
    total_daily_runoff = 0.;
    m->gw_recharge[i] = 0.;
    dailyCellLandAET = 0.;
    dailyCellLandAET_uncorr = 0.;
  } else {
    m->soil[i] *= lafPrev / landAreaFrac;
    initialStorage = m->soil[i];
    soil_water_overflow = 0;
    if (m->soil[i] > smax) {
      soil_water_overflow = m->soil[i] - smax;
      m->soil[i] = smax;
    }
    if (TempElevMax > T_SNOWFZ) {
      if (smax > 0.) {
        soil_saturation = m->soil[i] / smax;
        daily_runoff = dailyEffPrec * POW(soil_saturation, (double)m->gamma[i]);
        if (dailySoilPET > (maxDailyPET - dailyCanopyEvapo) * soil_saturation)
          dailyAET = (maxDailyPET - dailyCanopyEvapo) * soil_saturation;
        else
          dailyAET = dailySoilPET;
        m->soil[i] += dailyEffPrec - dailyAET - daily_runoff;
        if (fabs(m->soil[i]) <= MIN_STOR_VOL) m->soil[i] = 0.;
        dailyEffPrec = 0.;
        if (m->soil[i] < 0.) {
          dailyAET += m->soil[i];
          m->soil[i] = 0.;
        }
        daily_runoff *= cfa;
        immediate_runoff *= cfa;
        {
          const double rgmax = st->rgmax[i] / 100.;
          const float fg = st->gw_factor[i];
          if (((arid_gw) && (st->texture[i] < 21)) && (st->ldd[i] >= 0)) {
            pot_gw_recharge = 0.;
            if (rgmax < (fg * daily_runoff))
              daily_gw_recharge = rgmax;
            else
              daily_gw_recharge = fg * daily_runoff;
            if (dailyPrec <= P(WG_P_PCRITGWA)) {
              pot_gw_recharge = daily_gw_recharge;
              daily_gw_recharge = 0.;
            }
          } else {
            pot_gw_recharge = 0.;
            if (rgmax < (fg * daily_runoff))
              daily_gw_recharge = rgmax;
            else
              daily_gw_recharge = fg * daily_runoff;
          }
        }
        daily_runoff -= pot_gw_recharge;
        pot_gw_recharge /= cfa;
        m->soil[i] += pot_gw_recharge;
        if (m->soil[i] > smax) {
          soil_water_overflow += m->soil[i] - smax;
          m->soil[i] = smax;
        }
        soil_water_overflow *= cfa;
        total_daily_runoff = daily_runoff + immediate_runoff + soil_water_overflow;
      } else {
        total_daily_runoff = 0.;
        daily_gw_recharge = 0.;
      }
    } else {
      soil_water_overflow *= cfa;
      dailyEffPrec *= cfa;
      total_daily_runoff += soil_water_overflow + dailyEffPrec;
      daily_gw_recharge = 0.;
      dailyAET = 0.;
    }
    m->gw_recharge[i] = daily_gw_recharge;
    landStorageChangeSum += m->soil[i] - initialStorage;
    dailyCellLandAET = landStorageChangeSum * (cfa - 1.0) - dailyPrec * (cfa - 1.0) +
                       (dailyAET + dailyCanopyEvapo + dailySnowEvapo) * cfa;
    if (dailyCellLandAET < 0.) {
      neg_land_aet = dailyCellLandAET;
      dailyCellLandAET = 0.;
    }
    dailyCellLandAET_uncorr = (dailyAET + dailyCanopyEvapo + dailySnowEvapo);
  }

  /* daily.cpp:1241-1257 surface runoff */
  if (neg_land_aet < 0.) total_daily_runoff = total_daily_runoff + neg_land_aet;
  if (total_daily_runoff < 0.) total_daily_runoff = 0.;
  if ((total_daily_runoff - daily_gw_recharge) < 0.) {
    neg_runoff = total_daily_runoff - daily_gw_recharge;
    daily_gw_recharge = total_daily_runoff;
    m->soil[i] += neg_runoff;
    neg_runoff = 0.;
  }
  m->surf_runoff[i] = total_daily_runoff - daily_gw_recharge;
  m->aet_uncorr[i] = dailyCellLandAET_uncorr;
  (void)dailyCellLandAET;
  (void)snowmelt;
  (void)daily_snow_to_soil;
#undef P
}

/* ------------------------------------------------------------------ */
/* routing: routing.cpp:1629-5244                                       */
/* ------------------------------------------------------------------ */

/* routing.cpp:7309-7341 getNewRiverVelocity */
static double new_river_velocity(double slope, double bottom, double rough, double stor,
                                 double length, double mult) {
  double csa = (stor / length) * 1000000.;
  double depth = -bottom / 4.0 + sqrt(bottom * bottom / 16.0 + 0.5 * csa);
  double wp = bottom + 2.0 * depth * sqrt(5.0);
  double hr = csa / wp;
  double v = 1. / (mult * rough) * POW(hr, (2. / 3.)) * POW(slope, 0.5);
  v = v * 86.4;
  if (v < 0.00001)
    return 0.00001;
  else
    return v;
}

/* potCellRunoff's open-water term, routing.cpp:2047-2051: G_lakeBalance
 * (daily.cpp:773) in mm over the global lake, reservoir and wetland area. */
static double pot_open_water(const wgo_model *m, int32_t i) {
  const wg_static_fields *st = m->st;
  const double lakeBalance = (m->ow_prec[i] - m->ow_pet[i]) * m->cfa[i];
  return ((lakeBalance / 1000000.0) *
          (((double)st->lake_area[i] + (double)st->reservoir_area[i]) +
           ((st->cell_area[i] / 100.0) * (st->loc_lake[i] + st->loc_wetland[i] + st->glo_wetland[i]))));
}

/* One cell of the serial routing loop, routing.cpp:1838-4517 (parity profile). */
#define USE(f) (m->use[(size_t)(f) * (size_t)m->n + (size_t)i])

/* updateNetAbstractionGW, routing.cpp:5503-5572: the day's net abstraction from
 * groundwater, with the return flows of irrigation from surface water reduced
 * when yesterday's surface-water use went unsatisfied (G_dailyRemainingUse > 0)
 * or given back when a delayed use was satisfied (< 0). */
static double update_net_abstraction_gw(wgo_model *m, int32_t i, double nug) {
  double NAgnew = 0., returnflowChange = 0., eff = 0., frgi = 0., factor = 0., WUsiNew = 0.;
  double dailyRemainingUseFromIrrig = 0., reintroducedRatio = 0.;
  const double wu = USE(WG_USE_WITHDRAWAL_IRRIG), cu = USE(WG_USE_CONSUMPTIVE_IRRIG);
  if (USE(WG_USE_DAILY_REMAINING) > 1.e-12) {
    if (wu > 0.) {
      eff = cu / wu;
      frgi = m->use_frgi[i];
      factor = 1 - (1 - frgi) * (1 - eff);
      WUsiNew = 1 / factor * (wu * factor - USE(WG_USE_DAILY_REMAINING));
      if (WUsiNew < 0.) {
        WUsiNew = 0.;
        USE(WG_USE_UNSAT_IRRIG) += wu * factor;
        USE(WG_USE_UNSAT_OTHER) += USE(WG_USE_DAILY_REMAINING) - (wu * factor);
      } else {
        USE(WG_USE_UNSAT_IRRIG) += USE(WG_USE_DAILY_REMAINING);
      }
      returnflowChange = (frgi * (1 - eff) * (WUsiNew - wu));
      USE(WG_USE_REDUCED_RETURN) += returnflowChange;
      NAgnew = nug - returnflowChange;
      USE(WG_USE_DAILY_REMAINING) = 0.;
      return NAgnew;
    }
    USE(WG_USE_UNSAT_OTHER) += USE(WG_USE_DAILY_REMAINING);
    return nug;
  } else if (USE(WG_USE_DAILY_REMAINING) < -1.e-12) {
    dailyRemainingUseFromIrrig = USE(WG_USE_DAILY_REMAINING) + USE(WG_USE_UNSAT_OTHER);
    if (dailyRemainingUseFromIrrig < 0.) {
      USE(WG_USE_UNSAT_OTHER) = 0.;
    } else {
      USE(WG_USE_UNSAT_OTHER) += USE(WG_USE_DAILY_REMAINING);
      USE(WG_USE_DAILY_REMAINING) = 0.;
      return nug;
    }
    if (USE(WG_USE_UNSAT_IRRIG) == 0.) {
      USE(WG_USE_DAILY_REMAINING) = 0.;
      return nug;
    }
    reintroducedRatio = (dailyRemainingUseFromIrrig / USE(WG_USE_UNSAT_IRRIG));
    if (reintroducedRatio < -1.) {
      reintroducedRatio = -1.;
      dailyRemainingUseFromIrrig = USE(WG_USE_UNSAT_IRRIG) * -1.;
    }
    returnflowChange = (reintroducedRatio * USE(WG_USE_REDUCED_RETURN));
    USE(WG_USE_UNSAT_IRRIG) += dailyRemainingUseFromIrrig;
    USE(WG_USE_REDUCED_RETURN) += returnflowChange;
    NAgnew = nug - returnflowChange;
    USE(WG_USE_DAILY_REMAINING) = 0.;
    return NAgnew;
  }
  return nug;
}

/* G_dailydailyNUg of the day (routing.cpp:7440-7441), adapted when yesterday left
 * a remaining use (routing.cpp:1990-1994 and the other groundwater branches) */
static double net_abstraction_gw(wgo_model *m, int32_t i, int month) {
  double nug = m->use_nug[(size_t)i * 12 + month];
  if (USE(WG_USE_DAILY_REMAINING) != 0.) nug = update_net_abstraction_gw(m, i, nug);
  return nug;
}

/* routing.cpp:3590-3622, 3898-3919: what the river left of the use is taken from
 * the local lake (down to -maxStorage), whose area reduction factor is then
 * recomputed; the unsatisfied use is kept for the next day; the month's
 * irrigation withdrawal and consumptive use are kept for updateNetAbstractionGW. */
static void use_day_end(wgo_model *m, int32_t i, double remainingUse, int month) {
  const wg_static_fields *st = m->st;
  if ((st->loc_lake[i] > 0.) && (remainingUse > 0.)) {
    const double evarexp = st->params[(size_t)WG_M_EVAREDEX * m->n + i];
    const double maxStorage = ((st->loc_lake[i]) / 100.) * st->cell_area[i] * m->lake_depth[i];
    if (m->loc_lake[i] > (-1.) * maxStorage) {
      if (remainingUse < (maxStorage + m->loc_lake[i])) {
        m->loc_lake[i] -= remainingUse;
        remainingUse = 0.;
      } else {
        remainingUse -= (maxStorage + m->loc_lake[i]);
        m->loc_lake[i] = (-1.) * maxStorage;
      }
    }
    double r = 1. - POW(fabs(m->loc_lake[i] - maxStorage) / (2. * maxStorage), (evarexp * EVAPO_RED_EXP));
    if (r < 0.) r = 0.;
    if (r > 1.) r = 1.;
    m->red_ll[i] = r;
  }
  if (m->use_delayed == 1) {
    USE(WG_USE_DAILY_REMAINING) = remainingUse - USE(WG_USE_PREV_TOTAL_UNSAT);
    USE(WG_USE_TOTAL_UNSAT) = remainingUse;
  } else {
    USE(WG_USE_TOTAL_UNSAT) += remainingUse;
    USE(WG_USE_DAILY_REMAINING) = remainingUse;
  }
  USE(WG_USE_WITHDRAWAL_IRRIG) = m->use_wu[(size_t)i * 12 + month];
  USE(WG_USE_CONSUMPTIVE_IRRIG) = m->use_cu[(size_t)i * 12 + month];
}

static void route_cell(wgo_model *m, int32_t i, int day, int month) {
  const wg_static_fields *st = m->st;
  const int32_t n = m->n;
  const double *par = st->params;
#define P(id) (par[(size_t)(id) * n + i])
  const double cellArea = st->cell_area[i];
  const double cf = st->contfreq[i];
  const double k_gw = P(WG_P_GWOUTF_C);
  const double evarexp = P(WG_M_EVAREDEX);
  const double cfa = m->cfa[i];
  const int ldd = st->ldd[i];
  const int arid = st->aindex[i];
  double dailyLocalSurfaceRunoff, fswb_catchment = 0.;
  double localRunoff = 0., localRunoffIntoRiver = 0., localGWRunoffIntoRiver = 0.;
  double netGWin, gw_prev, gw_runoff;
  double inflow, totalInflow, outflow = 0., maxStorage;
  double gwr_loclak = 0., gwr_locwet = 0., gwr_glolak = 0., gwr_glowet = 0., gwr_res = 0.;

  if (!m->contcell[i]) return;

  /* routing.cpp:1885-1891 */
  if (m->laf[i] <= 0.)
    dailyLocalSurfaceRunoff = m->transfer[i] * cellArea / 1000000. * m->laf_prev[i] / 100.;
  else
    dailyLocalSurfaceRunoff = m->surf_runoff[i] * cellArea / 1000000. * m->laf[i] / 100.;

  /* routing.cpp:1898-1906 */
  if (ldd >= 0) {
    fswb_catchment = m->fswb_init[i] * 20.;
    if (fswb_catchment > 1.) fswb_catchment = 1.;
    localRunoffIntoRiver = (1. - fswb_catchment) * dailyLocalSurfaceRunoff;
  }
  if ((1 == arid) && (ldd >= 0)) {
    /* routing.cpp:1908-1916 (aridareaOpt=1: storage computed later) */
    localRunoff = fswb_catchment * dailyLocalSurfaceRunoff;
  }
  if ((0 == arid) && (ldd >= 0)) {
    /* routing.cpp:1968-2039 humid cells */
    gw_prev = m->gw[i];
    netGWin = (double)m->gw_recharge[i] * cellArea * ((double)m->laf[i] / 100.) / 1000000.;
    /* routing.cpp:1990-1996 */
    if (m->use_on) netGWin -= net_abstraction_gw(m, i, month);
    m->gw[i] = gw_prev * EXP(-1. * k_gw) + (1. / k_gw) * netGWin * (1. - EXP(-1. * k_gw));
    if (fabs(m->gw[i]) <= MIN_STOR_VOL) m->gw[i] = 0.;
    gw_runoff = gw_prev - m->gw[i] + netGWin;
    if (gw_runoff <= 0.) {
      gw_runoff = 0.;
      m->gw[i] = gw_prev + netGWin;
      if (fabs(m->gw[i]) <= MIN_STOR_VOL) m->gw[i] = 0.;
    }
    localGWRunoffIntoRiver = (1. - fswb_catchment) * gw_runoff;
    double localGWRunoff = fswb_catchment * gw_runoff;
    localRunoff = (fswb_catchment * dailyLocalSurfaceRunoff) + localGWRunoff;
    /* routing.cpp:2045-2053 */
    m->pot_acc[i] += localRunoff + localRunoffIntoRiver + localGWRunoffIntoRiver + pot_open_water(m, i);
  }
  if (ldd < 0) {
    /* routing.cpp:2112-2159 inland sinks */
    gw_prev = m->gw[i];
    netGWin = m->gw_recharge[i] * cellArea * (m->laf[i] / 100.) / 1000000.;
    /* routing.cpp:2131-2137 */
    if (m->use_on) netGWin -= net_abstraction_gw(m, i, month);
    m->gw[i] = gw_prev * EXP(-1. * k_gw) + (1. / k_gw) * netGWin * (1. - EXP(-1. * k_gw));
    if (fabs(m->gw[i]) <= MIN_STOR_VOL) m->gw[i] = 0.;
    gw_runoff = gw_prev - m->gw[i] + netGWin;
    if (gw_runoff <= 0.) {
      gw_runoff = 0.;
      m->gw[i] = gw_prev + netGWin;
      if (fabs(m->gw[i]) <= MIN_STOR_VOL) m->gw[i] = 0.;
    }
    if (m->laf[i] == 0.)
      dailyLocalSurfaceRunoff = m->transfer[i] * cellArea / 1000000. * m->laf_prev[i] / 100.;
    else
      dailyLocalSurfaceRunoff = m->surf_runoff[i] * cellArea / 1000000. * m->laf[i] / 100.;
    localRunoff = dailyLocalSurfaceRunoff + gw_runoff;
    /* routing.cpp:2180-2187 */
    m->pot_acc[i] += localRunoff + pot_open_water(m, i);
  }

  /* routing.cpp:2193-2296 (use_alloc = 0, aggrNUsGloLakResOpt = 0): the day's
   * desired surface-water use; remainingUse is 0 without water use */
  double remainingUse = 0.;
  if (m->use_on) {
    double totalDesiredUse = m->use_nus[(size_t)i * 12 + month];
    if (m->use_delayed == 1) {
      totalDesiredUse += USE(WG_USE_TOTAL_UNSAT);
      USE(WG_USE_PREV_TOTAL_UNSAT) = USE(WG_USE_TOTAL_UNSAT);
      USE(WG_USE_TOTAL_UNSAT) = 0;
    }
    remainingUse = totalDesiredUse;
  }
  double remainingUseGloLake = 0., remainingUseRes = 0.;

  /* routing.cpp:2308-2313 */
  inflow = localRunoff;
  const double k_sw = P(WG_P_SWOUTF_C);
  const double ow_prec = m->ow_prec[i], ow_pet = m->ow_pet[i];
  const int arid_swb = (1 == arid) && (ldd >= 0);

  /* routing.cpp:2316-2490 local lake */
  if (st->loc_lake[i] > 0.) {
    double prev = m->loc_lake[i], rf = m->red_ll[i], evap, PETgwr, PETgwrMax;
    maxStorage = ((st->loc_lake[i]) / 100.) * cellArea * m->lake_depth[i];
    evap = ((1.0 - cfa) * ow_prec * rf) + (cfa * (ow_pet * rf));
    if (evap < 0.) evap = 0.;
    totalInflow = inflow + (ow_prec * rf) * (cellArea / 1000000.) * (st->loc_lake[i] / 100.);
    if (arid_swb)
      gwr_loclak = KSWBGW * rf * st->loc_lake[i] / 100. / (cf / 100.);
    else
      gwr_loclak = 0.;
    PETgwr = evap * (cellArea / 1000000.) * (st->loc_lake[i] / 100.) +
             gwr_loclak * cellArea * (cf / 100.) / 1000000.;
    PETgwrMax = prev + maxStorage + totalInflow;
    if (PETgwrMax < 0.) PETgwrMax = 0.;
    if (PETgwr > PETgwrMax) {
      m->loc_lake[i] = (-1.) * maxStorage;
      gwr_loclak *= PETgwrMax / PETgwr;
    } else {
      m->loc_lake[i] = prev + totalInflow - PETgwr;
    }
    if (prev > 0.) {
      outflow = k_sw * prev * POW((prev / maxStorage), LAKE_OUTFLOW_EXP);
      if (m->loc_lake[i] <= 0.) {
        outflow = 0;
      } else {
        if (outflow > m->loc_lake[i]) outflow = m->loc_lake[i];
      }
    } else {
      outflow = 0.;
    }
    m->loc_lake[i] -= outflow;
    if (fabs(m->loc_lake[i]) <= MIN_STOR_VOL) m->loc_lake[i] = 0.;
    if (m->loc_lake[i] > maxStorage) {
      outflow += (m->loc_lake[i] - maxStorage);
      m->loc_lake[i] = maxStorage;
    }
    inflow = outflow;
    m->red_ll[i] = 1. - POW(fabs(m->loc_lake[i] - maxStorage) / (2. * maxStorage), (evarexp * EVAPO_RED_EXP));
    if (m->red_ll[i] < 0.) m->red_ll[i] = 0.;
    if (m->red_ll[i] > 1.) m->red_ll[i] = 1.;
  }

  /* routing.cpp:2495-2617 local wetland */
  if (st->loc_wetland[i] > 0.) {
    double prev = m->loc_wetl[i], rf = m->red_lw[i], evap, PETgwr, PETgwrMax;
    maxStorage = ((st->loc_wetland[i]) / 100.) * cellArea * m->wetl_depth[i];
    evap = ((1.0 - cfa) * ow_prec * rf) + (cfa * (ow_pet * rf));
    if (evap < 0.) evap = 0.;
    totalInflow = inflow + (ow_prec * rf * (cellArea / 1000000.) * (st->loc_wetland[i] / 100.));
    if (arid_swb)
      gwr_locwet = KSWBGW * rf * st->loc_wetland[i] / 100. / (cf / 100.);
    else
      gwr_locwet = 0.;
    PETgwr = evap * (cellArea / 1000000.) * (st->loc_wetland[i] / 100.) +
             gwr_locwet * cellArea * (cf / 100.) / 1000000.;
    PETgwrMax = prev + totalInflow;
    if (PETgwr > PETgwrMax) {
      m->loc_wetl[i] = 0.;
      gwr_locwet *= PETgwrMax / PETgwr;
    } else {
      m->loc_wetl[i] = prev + totalInflow - PETgwr;
    }
    if (fabs(m->loc_wetl[i]) <= MIN_STOR_VOL) m->loc_wetl[i] = 0.;
    if (m->loc_wetl[i] > 0.) {
      outflow = k_sw * m->loc_wetl[i] * POW((m->loc_wetl[i] / maxStorage), WETL_OUTFLOW_EXP);
      if (outflow > m->loc_wetl[i]) outflow = m->loc_wetl[i];
    } else {
      outflow = 0.;
    }
    m->loc_wetl[i] -= outflow;
    if (m->loc_wetl[i] > maxStorage) {
      outflow += (m->loc_wetl[i] - maxStorage);
      m->loc_wetl[i] = maxStorage;
    }
    inflow = outflow;
    m->red_lw[i] = 1. - POW(fabs(m->loc_wetl[i] - maxStorage) / (maxStorage), (evarexp * EVAPO_RED_EXP));
    if (m->red_lw[i] < 0.) m->red_lw[i] = 0.;
    if (m->red_lw[i] > 1.) m->red_lw[i] = 1.;
  }

  /* routing.cpp:2623-2624: the only read of upstream water */
  inflow += m->river_inflow[i];

  /* routing.cpp:2630-2804 global lake */
  if (st->lake_area[i] > 0.) {
    double prev = m->glo_lake[i], rf = m->red_gl[i], evap, use, useMax;
    maxStorage = ((double)st->lake_area[i]) * m->lake_depth[i];
    evap = ((1.0 - cfa) * ow_prec) + (cfa * ow_pet * rf);
    if (evap < 0.) evap = 0.;
    totalInflow = inflow + (ow_prec * ((double)st->lake_area[i] / 1000000.));
    if (arid_swb)
      gwr_glolak = KSWBGW * rf * ((double)st->lake_area[i] / (cellArea * (cf / 100.)));
    else
      gwr_glolak = 0.;
    /* routing.cpp:2680-2689: half the use when the cell also has a reservoir */
    if (st->reservoir_area[i] > 0.)
      remainingUseGloLake = 0.5 * remainingUse;
    else
      remainingUseGloLake = remainingUse;
    use = evap * ((double)st->lake_area[i] / 1000000.) + gwr_glolak * cellArea * (cf / 100.) / 1000000. +
          remainingUseGloLake;
    useMax = totalInflow + maxStorage + prev;
    if (use > useMax) {
      m->glo_lake[i] = (-1.) * maxStorage;
      outflow = 0.;
      gwr_glolak *= useMax / use;
      /* routing.cpp:2710-2716 */
      if (remainingUseGloLake > 0.)
        remainingUseGloLake -= remainingUseGloLake * useMax / use;
      else
        remainingUseGloLake = 0.;
    } else {
      m->glo_lake[i] = prev * EXP(-1. * k_sw) + (1. / k_sw) * (totalInflow - use) * (1. - EXP(-1. * k_sw));
      outflow = totalInflow + prev - m->glo_lake[i] - use;
      if (m->glo_lake[i] > maxStorage) {
        outflow += (m->glo_lake[i] - maxStorage);
        m->glo_lake[i] = maxStorage;
      }
      if (outflow < 0.) {
        outflow = 0.;
        m->glo_lake[i] = prev + totalInflow - use;
      }
      remainingUseGloLake = 0.; /* routing.cpp:2752 */
    }
    if (fabs(m->glo_lake[i]) <= MIN_STOR_VOL) m->glo_lake[i] = 0.;
    inflow = outflow;
    m->red_gl[i] = 1. - POW(fabs(m->glo_lake[i] - maxStorage) / (2. * maxStorage), (evarexp * EVAPO_RED_EXP));
    if (m->red_gl[i] < 0.) m->red_gl[i] = 0.;
    if (m->red_gl[i] > 1.) m->red_gl[i] = 1.;
  }

  /* routing.cpp:2807-3082 reservoir */
  if (st->reservoir_area[i] > 0.) {
    double prev = m->res[i], rf = m->red_res[i], evap, PETgwr, PETgwrMax;
    double c_ratio, prov_rel = 0., release;
    const double cap = st->stor_cap[i], qmean = st->mean_outflow[i];
    c_ratio = cap / (qmean * 31536000. / 1000000000.);
    maxStorage = cap;
    evap = ((1.0 - cfa) * ow_prec) + (cfa * (ow_pet * rf));
    if (evap < 0.) evap = 0.;
    totalInflow = inflow + (ow_prec * ((double)st->reservoir_area[i] / 1000000.));
    if (arid_swb)
      gwr_res = KSWBGW * rf * ((double)st->reservoir_area[i] / (cellArea * (cf / 100.)));
    else
      gwr_res = 0.;
    PETgwr = evap * ((double)st->reservoir_area[i] / 1000000.) + gwr_res * cellArea * (cf / 100.) / 1000000.;
    PETgwrMax = prev + totalInflow;
    if (PETgwr > PETgwrMax) {
      m->res[i] = prev + totalInflow - PETgwrMax;
      gwr_res *= PETgwrMax / PETgwr;
    } else {
      m->res[i] = prev + totalInflow - PETgwr;
    }
    /* routing.cpp:2869-2874, 2922-2945: the use down to 10 % of the capacity;
     * a negative use (return flow) is added */
    if (st->lake_area[i] > 0.)
      remainingUseRes = 0.5 * remainingUse + remainingUseGloLake;
    else
      remainingUseRes = remainingUse;
    if (m->use_on) {
      if (remainingUseRes < 0.) {
        m->res[i] -= remainingUseRes;
        remainingUseRes = 0.;
      } else {
        if (m->res[i] > (0.1 * cap)) {
          if (remainingUseRes < (m->res[i] - (cap * 0.1))) {
            m->res[i] -= remainingUseRes;
            remainingUseRes = 0.;
          } else {
            remainingUseRes -= (m->res[i] - (cap * 0.1));
            m->res[i] = cap * 0.1;
          }
        }
      }
    }
    if (fabs(m->res[i]) <= MIN_STOR_VOL) m->res[i] = 0.;
    if (month == st->start_month[i] - 1) {
      if (day == first_day_in_month_r[month]) {
        if (m->res[i] < (cap * 0.1))
          m->k_release[i] = 0.1;
        else
          m->k_release[i] = m->res[i] / (maxStorage * 0.85);
      }
    }
    if ((st->res_type[i] + 0) == 1) {
      /* irrigation reservoir: the daily net use of the cell and of up to five
       * downstream cells (routing.cpp:2972-2983; 0 without water use) */
      const double mean_demand = m->mean_demand[i];
      double dailyUse = 0.0;
      if (m->use_on) {
        dailyUse = m->use_nus[(size_t)i * 12 + month];
        short k = 0;
        int32_t downstreamCell = st->downstream[i];
        while (k < 5 && downstreamCell > 0) {
          /* G_reservoir_area[downstreamCell], the cell after the downstream cell (routing.cpp:2981) */
          const double area_next = downstreamCell < n ? st->reservoir_area[downstreamCell] : 0.;
          if (!(area_next <= 0)) break;
          const double coeff = st->alloc_coeff ? st->alloc_coeff[(size_t)i * 5 + k] : 1.0;
          k++;
          dailyUse += m->use_nus[(size_t)(downstreamCell - 1) * 12 + month] * coeff;
          downstreamCell = st->downstream[downstreamCell - 1];
        }
      }
      double monthlyUse = dailyUse * days_in_month[month];
      monthlyUse = monthlyUse * 1000000000. / (days_in_month[month] * 86400.);
      if (mean_demand >= 0.5 * qmean)
        prov_rel = qmean / 2. * (1. + monthlyUse / mean_demand);
      else
        prov_rel = qmean + monthlyUse - mean_demand;
    } else if ((st->res_type[i] + 0) == 2) {
      prov_rel = qmean;
    }
    if (c_ratio >= 0.5)
      release = m->k_release[i] * prov_rel;
    else
      release = ((4. * c_ratio * c_ratio) * m->k_release[i] * prov_rel) +
                ((1.0 - ((4. * c_ratio * c_ratio))) * inflow * 1000000000. / (24. * 3600.));
    if (m->res[i] >= (cap * 0.1))
      outflow = release * (24. * 3600.) / 1000000000.;
    else
      outflow = 0.1 * release * (24. * 3600.) / 1000000000.;
    if (outflow < 0.) outflow = 0.;
    m->res[i] -= outflow;
    if (m->res[i] > maxStorage) {
      outflow += (m->res[i] - maxStorage);
      m->res[i] = maxStorage;
    }
    if (m->res[i] < 0.) {
      outflow += m->res[i];
      m->res[i] = 0.;
    }
    inflow = outflow;
    m->red_res[i] = 1. - POW(fabs(m->res[i] - maxStorage) / maxStorage, EVAPO_RED_EXP_RES);
    if (m->red_res[i] < 0.) m->red_res[i] = 0.;
    if (m->red_res[i] > 1.) m->red_res[i] = 1.;
  }

  /* routing.cpp:3083-3160: with aggrNUsGloLakResOpt = 1 every global lake or
   * reservoir cell scans all cells for riparian cells of its lake unit; without
   * water use nothing is assigned.  The scan's cost is kept for the baseline's
   * footnote (each cell is its own unit here, so the scan finds only itself). */
  if (m->aggr_scan && (st->reservoir_area[i] > 0. || st->lake_area[i] > 0.)) {
    int64_t hits = 0;
    for (int32_t k = 0; k < n; k++) hits += (m->routing_cell[k] == i + 1);
    m->aggr_matches += hits;
  }

  /* routing.cpp:3162-3171 (aggrNUsGloLakResOpt = 0) */
  if (st->reservoir_area[i] > 0. || st->lake_area[i] > 0.) {
    if (st->reservoir_area[i] > 0.)
      remainingUse = remainingUseRes;
    else
      remainingUse = remainingUseGloLake;
  }

  /* routing.cpp:3178-3297 global wetland */
  if (st->glo_wetland[i] > 0) {
    double prev = m->glo_wetl[i], rf = m->red_gw[i], evap, PETgwr, PETgwrMax;
    maxStorage = ((st->glo_wetland[i]) / 100.) * cellArea * m->wetl_depth[i];
    evap = ((1.0 - cfa) * (ow_prec * rf)) + (cfa * (ow_pet * rf));
    if (evap < 0.) evap = 0.;
    totalInflow = inflow + (ow_prec * rf * (cellArea / 1000000.) * (st->glo_wetland[i] / 100.));
    if (arid_swb)
      gwr_glowet = KSWBGW * rf * st->glo_wetland[i] / 100. / (cf / 100.);
    else
      gwr_glowet = 0.;
    PETgwr = evap * (cellArea / 1000000.) * ((st->glo_wetland[i]) / 100.) +
             gwr_glowet * cellArea * (cf / 100.) / 1000000.;
    PETgwrMax = totalInflow + prev;
    if (PETgwr > PETgwrMax) {
      m->glo_wetl[i] = 0.;
      outflow = 0.;
      gwr_glowet *= PETgwrMax / PETgwr;
    } else {
      m->glo_wetl[i] = prev * EXP(-1. * k_sw) + (1. / k_sw) * (totalInflow - PETgwr) * (1. - EXP(-1. * k_sw));
      outflow = totalInflow + prev - m->glo_wetl[i] - PETgwr;
    }
    if (m->glo_wetl[i] > maxStorage) {
      outflow += (m->glo_wetl[i] - maxStorage);
      m->glo_wetl[i] = maxStorage;
    }
    if (fabs(m->glo_wetl[i]) <= MIN_STOR_VOL) m->glo_wetl[i] = 0.;
    inflow = outflow;
    m->red_gw[i] = 1. - POW(fabs(m->glo_wetl[i] - maxStorage) / maxStorage, (evarexp * EVAPO_RED_EXP));
    if (m->red_gw[i] < 0.) m->red_gw[i] = 0.;
    if (m->red_gw[i] > 1.) m->red_gw[i] = 1.;
  }

  /* routing.cpp:3305-3386 groundwater of (semi-)arid cells, with recharge below SWB */
  if (arid_swb) {
    double gwr_swb = gwr_loclak + gwr_glolak + gwr_locwet + gwr_glowet + gwr_res;
    gw_prev = m->gw[i];
    netGWin = gwr_swb * cellArea * (cf / 100.) / 1000000. +
              (double)m->gw_recharge[i] * cellArea * ((double)m->laf[i] / 100.) / 1000000.;
    /* routing.cpp:3324-3330 */
    if (m->use_on) netGWin -= net_abstraction_gw(m, i, month);
    m->gw[i] = gw_prev * EXP(-1. * k_gw) + (1. / k_gw) * netGWin * (1. - EXP(-1. * k_gw));
    if (fabs(m->gw[i]) <= MIN_STOR_VOL) m->gw[i] = 0.;
    gw_runoff = gw_prev - m->gw[i] + netGWin;
    if (gw_runoff <= 0.) {
      gw_runoff = 0.;
      m->gw[i] = gw_prev + netGWin;
      if (fabs(m->gw[i]) <= MIN_STOR_VOL) m->gw[i] = 0.;
    }
    localGWRunoffIntoRiver = gw_runoff;
    /* routing.cpp:3372-3379 */
    m->pot_acc[i] += localRunoff + localRunoffIntoRiver + localGWRunoffIntoRiver + pot_open_water(m, i);
  }

  /* routing.cpp:3402-3413 */
  double riverInflow = inflow;
  if (ldd >= 0) {
    riverInflow += localRunoffIntoRiver;
    riverInflow += localGWRunoffIntoRiver;
  }
  /* routing.cpp:3424-3432 */
  const double L = m->river_length[i];
  double velocity = new_river_velocity(m->river_slope[i], m->bottom_width[i], m->roughness[i],
                                       m->river[i], L, P(WG_M_RIVRGH_C));
  double K = velocity / L;
  double prev = m->river[i];
  /* routing.cpp:3441-3460 */
  double riverAreaFrac = m->river_frac_next[i];
  double riverEvapo = ((1.0 - cfa) * (ow_prec) + (cfa * ow_pet)) * riverAreaFrac / 100. * cellArea / 1000000.;
  double riverPrecip = ow_prec * riverAreaFrac / 100. * cellArea / 1000000.;
  riverInflow += riverPrecip;
  double evapRemUse = remainingUse + riverEvapo;
  double evapRemUseMax = riverInflow + (K * prev * EXP(-1. * K)) / (1. - EXP(-1. * K));
  double tv;
  if (evapRemUse > evapRemUseMax) {
    m->river[i] = 0.;
    tv = riverInflow + prev - evapRemUseMax;
    if (tv < 0.) tv = 0.;
    /* routing.cpp:3489-3492 */
    if (remainingUse > 0.)
      remainingUse -= remainingUse * evapRemUseMax / evapRemUse;
    else
      remainingUse = 0.;
  } else {
    m->river[i] = prev * EXP(-1. * K) + (1. / K) * (riverInflow - evapRemUse) * (1. - EXP(-1. * K));
    if (fabs(m->river[i]) <= MIN_STOR_VOL) m->river[i] = 0.;
    tv = riverInflow + prev - m->river[i] - evapRemUse;
    if (tv < 0.) tv = 0.;
    remainingUse = 0.; /* routing.cpp:3512 */
  }
  /* routing.cpp:3546-3586 river width and area fraction for the next day */
  {
    const double B = m->bottom_width[i];
    double csa = m->river[i] / L;
    double depth = -B / (4. * 1000.) + sqrt(B / 1000. * B / (16. * 1000.) + 0.5 * csa);
    double width = B / 1000. + 4. * depth;
    if (width > m->width_bf[i] / 1000.) width = m->width_bf[i] / 1000.;
    m->red_river[i] = 1. - POW(fabs(m->river[i] - m->river_stor_max[i]) / m->river_stor_max[i],
                               (evarexp * EVAPO_RED_EXP));
    if (m->red_river[i] < 0.) m->red_river[i] = 0.;
    if (m->red_river[i] > 1.) m->red_river[i] = 1.;
    if (m->contcell[i])
      m->river_frac_next[i] = m->red_river[i] * L * width * 100. / cellArea;
    else
      m->river_frac_next[i] = 0.;
    m->river_frac_change[i] = m->river_frac_next[i] - riverAreaFrac;
  }
  if (m->use_on) use_day_end(m, i, remainingUse, month);
  m->discharge[i] = tv;
  /* routing.cpp:3955-3958 the only write to another cell */
  if ((st->downstream[i] - 1) >= 0) m->river_inflow[st->downstream[i] - 1] += tv;
  /* routing.cpp:4232-4238 station discharge: the first station of the cell; the
   * reference finds it by scanning all stations (kept as an option for timing) */
  if (m->station_scan) {
    for (int32_t b = 0; b < m->n_st; b++) {
      if (i == m->st_cell[b]) {
        if (ldd >= 0) m->st_sum[b] += tv;
        break;
      }
    }
  } else if (m->st_slot && m->st_slot[i] >= 0) {
    if (ldd >= 0) m->st_sum[m->st_slot[i]] += tv;
  }
  /* routing.cpp:4517 */
  m->river_inflow[i] = 0.;
#undef P
}

/* routing.cpp:5034-5188 fractions for the next day, and 5343-5352 updateLandAreaFrac */
static void fractions_cell(wgo_model *m, int32_t i) {
  const wg_static_fields *st = m->st;
  double fLocLake, fLocWet, fGloWet, fswb_old, change_pct, max_raf;
  if ((st->loc_lake[i] > 0.) && (m->red_ll[i] > 0.))
    fLocLake = (m->red_ll[i] * st->loc_lake[i] / 100.);
  else
    fLocLake = 0.;
  if ((st->loc_wetland[i] > 0.) && (m->red_lw[i] > 0.))
    fLocWet = (m->red_lw[i] * st->loc_wetland[i] / 100.);
  else
    fLocWet = 0.;
  if ((st->glo_wetland[i] > 0.) && (m->red_gw[i] > 0.))
    fGloWet = (m->red_gw[i] * st->glo_wetland[i] / 100.);
  else
    fGloWet = 0.;
  fswb_old = m->fswb_next[i];
  m->fswb_next[i] = fLocLake + fLocWet + fGloWet;
  change_pct = m->fswb_next[i] * 100. - fswb_old * 100.0;
  max_raf = st->contfreq[i] / 100. - m->f_glo_lake[i];
  if (((st->lake_area[i] > 0.) || (st->reservoir_area[i] > 0.)) && (m->f_glo_lake[i] == 1.)) {
    m->river_frac_next[i] = 0.;
    m->river_frac_change[i] = 0.;
    m->red_river[i] = 0.;
  } else {
    if (m->river_frac_next[i] <= max_raf) {
      if (m->fswb_next[i] > (max_raf - m->river_frac_next[i])) {
        double corr = (max_raf - m->river_frac_next[i]) / m->fswb_next[i];
        m->red_ll[i] *= corr;
        m->red_lw[i] *= corr;
        m->red_gw[i] *= corr;
        if ((fLocLake > 0.) && (m->red_ll[i] > 0.)) {
          fLocLake = (m->red_ll[i] * st->loc_lake[i] / 100.);
        } else {
          m->red_ll[i] = 0.;
          fLocLake = 0.;
        }
        if ((fLocWet > 0.) && (m->red_lw[i] > 0.)) {
          fLocWet = (m->red_lw[i] * st->loc_wetland[i] / 100.);
        } else {
          m->red_lw[i] = 0.;
          fLocWet = 0.;
        }
        if ((fGloWet > 0.) && (m->red_gw[i] > 0.)) {
          fGloWet = (m->red_gw[i] * st->glo_wetland[i] / 100.);
        } else {
          m->red_gw[i] = 0.;
          fGloWet = 0.;
        }
      }
    } else {
      double deficit = m->river_frac_next[i] - max_raf;
      m->river_frac_change[i] -= deficit;
      m->red_river[i] *= max_raf / m->river_frac_next[i];
      m->river_frac_next[i] = max_raf;
      fLocLake = 0.;
      fLocWet = 0.;
      fGloWet = 0.;
    }
  }
  m->fswb_next[i] = fLocLake + fLocWet + fGloWet;
  change_pct = m->fswb_next[i] * 100. - fswb_old * 100.;
  m->laf_next[i] = m->laf[i] - (change_pct + (m->river_frac_change[i] * 100.));
  if (m->laf_next[i] < 0.) m->laf_next[i] = 0.;
  /* routing.cpp:5343-5352 */
  m->laf_prev[i] = m->laf[i];
  m->laf[i] = m->laf_next[i];
}

void wgo_step_day(wgo_model *m, const float *temp, const float *prec, const float *sw, const float *lw,
                  int day, int month) {
  const int32_t n = m->n;
  /* integrateWGHM.cpp:767-783: calcNewDay for every continental cell; cells are
   * independent, so the OpenMP split over cells (as the reference's) changes no bit */
#pragma omp parallel for schedule(static) num_threads(m->threads)
  for (int32_t i = 0; i < n; i++)
    if (m->contcell[i]) vertical_cell(m, i, temp[i], prec[i], sw[i], lw[i]);
  /* routing.cpp:1775-1811: zero the inflow accumulators */
  for (int32_t i = 0; i < n; i++) {
    m->river_inflow[i] = 0.;
    m->discharge[i] = 0.;
  }
  /* routing.cpp:1835-1838: serial loop in routing order */
  for (int32_t rc = 0; rc < n; rc++) route_cell(m, m->routing_cell[rc] - 1, day, month);
  for (int32_t i = 0; i < n; i++) fractions_cell(m, i);
  m->started = 1;
}

void wgo_set_threads(wgo_model *m, int threads) { m->threads = threads > 0 ? threads : 1; }

void wgo_run(wgo_model *m, const float *forcing, int32_t n_days, int32_t first_doy) {
  const size_t n = (size_t)m->n;
  int doy = first_doy;
  for (int32_t d = 0; d < n_days; d++) {
    const float *f = forcing + (size_t)d * 4 * n;
    wgo_step_day(m, f, f + n, f + 2 * n, f + 3 * n, doy, wgo_month_of_day(doy));
    doy = doy % 365 + 1;
  }
}

/* ------------------------------------------------------------------ */
/* state access                                                         */
/* ------------------------------------------------------------------ */

#define CP_OUT(dst, src, cnt)                                  \
  do {                                                         \
    if (dst) memcpy((dst), (src), (size_t)(cnt) * sizeof(*(src))); \
  } while (0)
#define CP_IN(dst, src, cnt)                                   \
  do {                                                         \
    if (src) memcpy((dst), (src), (size_t)(cnt) * sizeof(*(dst))); \
  } while (0)

int wgo_get_state(const wgo_model *m, wg_state *o) {
  const int32_t n = m->n;
  CP_OUT(o->canopy, m->canopy, n); CP_OUT(o->soil, m->soil, n); CP_OUT(o->snow, m->snow, n);
  CP_OUT(o->snow_elev, m->snow_elev, (size_t)n * WG_NELEV); CP_OUT(o->storage_transfer, m->transfer, n);
  CP_OUT(o->lai_days, m->lai_days, n); CP_OUT(o->lai_status, m->lai_status, n);
  CP_OUT(o->lai_precsum, m->lai_precsum, n);
  CP_OUT(o->groundwater, m->gw, n); CP_OUT(o->loc_lake, m->loc_lake, n); CP_OUT(o->loc_wetl, m->loc_wetl, n);
  CP_OUT(o->glo_lake, m->glo_lake, n); CP_OUT(o->glo_wetl, m->glo_wetl, n); CP_OUT(o->reservoir, m->res, n);
  CP_OUT(o->river, m->river, n);
  CP_OUT(o->red_loc_lake, m->red_ll, n); CP_OUT(o->red_loc_wetl, m->red_lw, n);
  CP_OUT(o->red_glo_lake, m->red_gl, n); CP_OUT(o->red_glo_wetl, m->red_gw, n);
  CP_OUT(o->red_res, m->red_res, n); CP_OUT(o->red_river, m->red_river, n);
  CP_OUT(o->river_frac_next, m->river_frac_next, n); CP_OUT(o->fswb_frac_next, m->fswb_next, n);
  CP_OUT(o->laf, m->laf, n); CP_OUT(o->laf_prev, m->laf_prev, n); CP_OUT(o->k_release, m->k_release, n);
  o->started = m->started;
  return 0;
}

int wgo_set_state(wgo_model *m, const wg_state *o) {
  const int32_t n = m->n;
  CP_IN(m->canopy, o->canopy, n); CP_IN(m->soil, o->soil, n); CP_IN(m->snow, o->snow, n);
  CP_IN(m->snow_elev, o->snow_elev, (size_t)n * WG_NELEV); CP_IN(m->transfer, o->storage_transfer, n);
  CP_IN(m->lai_days, o->lai_days, n); CP_IN(m->lai_status, o->lai_status, n);
  CP_IN(m->lai_precsum, o->lai_precsum, n);
  CP_IN(m->gw, o->groundwater, n); CP_IN(m->loc_lake, o->loc_lake, n); CP_IN(m->loc_wetl, o->loc_wetl, n);
  CP_IN(m->glo_lake, o->glo_lake, n); CP_IN(m->glo_wetl, o->glo_wetl, n); CP_IN(m->res, o->reservoir, n);
  CP_IN(m->river, o->river, n);
  CP_IN(m->red_ll, o->red_loc_lake, n); CP_IN(m->red_lw, o->red_loc_wetl, n);
  CP_IN(m->red_gl, o->red_glo_lake, n); CP_IN(m->red_gw, o->red_glo_wetl, n);
  CP_IN(m->red_res, o->red_res, n); CP_IN(m->red_river, o->red_river, n);
  CP_IN(m->river_frac_next, o->river_frac_next, n); CP_IN(m->fswb_next, o->fswb_frac_next, n);
  CP_IN(m->laf, o->laf, n); CP_IN(m->laf_prev, o->laf_prev, n); CP_IN(m->k_release, o->k_release, n);
  m->started = o->started;
  return 0;
}

int wgo_get_day_outputs(const wgo_model *m, wg_day_outputs *o) {
  const int32_t n = m->n;
  CP_OUT(o->discharge, m->discharge, n);
  CP_OUT(o->surface_runoff, m->surf_runoff, n);
  CP_OUT(o->gw_recharge, m->gw_recharge, n);
  CP_OUT(o->open_water_pet, m->ow_pet, n);
  CP_OUT(o->cell_aet_uncorr, m->aet_uncorr, n);
  CP_OUT(o->prec, m->prec_out, n);
  return 0;
}

void wgo_set_stations(wgo_model *m, int32_t n_st, const int32_t *cells) {
  free(m->st_cell);
  free(m->st_sum);
  free(m->st_slot);
  m->n_st = n_st;
  m->st_cell = (int32_t *)malloc((size_t)(n_st > 0 ? n_st : 1) * sizeof(int32_t));
  m->st_sum = (double *)calloc((size_t)(n_st > 0 ? n_st : 1), sizeof(double));
  m->st_slot = (int32_t *)malloc((size_t)m->n * sizeof(int32_t));
  for (int32_t i = 0; i < m->n; i++) m->st_slot[i] = -1;
  for (int32_t b = 0; b < n_st; b++) {
    m->st_cell[b] = cells[b];
    if (cells[b] >= 0 && cells[b] < m->n && m->st_slot[cells[b]] < 0) m->st_slot[cells[b]] = b;
  }
}

void wgo_set_reference_scans(wgo_model *m, int station_scan, int aggr_scan) {
  m->station_scan = station_scan;
  m->aggr_scan = aggr_scan;
}

void wgo_station_discharge(const wgo_model *m, double *sum) {
  for (int32_t b = 0; b < m->n_st; b++) sum[b] = m->st_sum[b];
}

void wgo_reset_station_discharge(wgo_model *m) {
  for (int32_t b = 0; b < m->n_st; b++) m->st_sum[b] = 0.;
}

void wgo_pot_cell_runoff(const wgo_model *m, double *pot) {
  for (int32_t i = 0; i < m->n; i++) pot[i] = m->pot_acc[i];
}

void wgo_reset_pot_cell_runoff(wgo_model *m) {
  for (int32_t i = 0; i < m->n; i++) m->pot_acc[i] = 0.;
}

/* ------------------------------------------------------------------ */
/* water use                                                            */
/* ------------------------------------------------------------------ */

int wgo_set_water_use(wgo_model *m, int delayed, const double *frac_return_gw) {
  const size_t n = (size_t)m->n;
  if (!m->use) {
    m->use = (double *)calloc(n * WG_USE_NFIELDS, sizeof(double));
    m->use_nus = (double *)calloc(n * 12, sizeof(double));
    m->use_nug = (double *)calloc(n * 12, sizeof(double));
    m->use_wu = (double *)calloc(n * 12, sizeof(double));
    m->use_cu = (double *)calloc(n * 12, sizeof(double));
    m->use_frgi = (double *)calloc(n, sizeof(double));
    if (!m->use || !m->use_nus || !m->use_nug || !m->use_wu || !m->use_cu || !m->use_frgi) return -4;
  }
  memcpy(m->use_frgi, frac_return_gw, n * sizeof(double));
  m->use_on = 1;
  m->use_delayed = delayed;
  return 0;
}

/* dailyNUInit, routing.cpp:909-972: the year's [N][12] tables in m3 per month;
 * NUs and NUg times their multipliers, then over 1e9 x days of the month; the
 * irrigation withdrawal and consumptive use over the same (routing.cpp:3918-3919) */
void wgo_water_use_year(wgo_model *m, const double *nus_m3, const double *nug_m3, const double *wu_m3,
                        const double *cu_m3) {
  const int32_t n = m->n;
  const double *par = m->st->params;
  for (int32_t i = 0; i < n; i++) {
    for (int month = 0; month < 12; month++) {
      const size_t k = (size_t)i * 12 + month;
      double nus = par[(size_t)WG_M_NETABSSW * n + i] * nus_m3[k];
      double nug = par[(size_t)WG_M_NETABSGW * n + i] * nug_m3[k];
      m->use_nus[k] = nus / (1000000000. * (double)days_in_month[month]);
      m->use_nug[k] = nug / (1000000000. * (double)days_in_month[month]);
      m->use_wu[k] = wu_m3[k] / (1000000000. * (double)days_in_month[month]);
      m->use_cu[k] = cu_m3[k] / (1000000000. * (double)days_in_month[month]);
    }
  }
}

/* annualWaterUsePostProcessing, routing.cpp:5257-5296 */
void wgo_water_use_year_end(wgo_model *m) {
  if (!m->use_on) return;
  for (int32_t i = 0; i < m->n; i++) {
    if (USE(WG_USE_TOTAL_UNSAT) < USE(WG_USE_UNSAT_PREV_YEAR)) {
      USE(WG_USE_UNSAT_PREV_YEAR) = USE(WG_USE_TOTAL_UNSAT);
      USE(WG_USE_UNSAT_IRRIG_PREV_YEAR) = USE(WG_USE_UNSAT_IRRIG);
      USE(WG_USE_UNSAT_OTHER_PREV_YEAR) = USE(WG_USE_UNSAT_OTHER);
      USE(WG_USE_REDUCED_RETURN_PREV_YEAR) = USE(WG_USE_REDUCED_RETURN);
      USE(WG_USE_UNSAT_IRRIG) = 0.;
      USE(WG_USE_UNSAT_OTHER) = 0.;
      USE(WG_USE_REDUCED_RETURN) = 0.;
      USE(WG_USE_TOTAL_UNSAT) = 0.;
    } else {
      USE(WG_USE_TOTAL_UNSAT) -= USE(WG_USE_UNSAT_PREV_YEAR);
      USE(WG_USE_UNSAT_IRRIG) -= USE(WG_USE_UNSAT_IRRIG_PREV_YEAR);
      USE(WG_USE_UNSAT_OTHER) -= USE(WG_USE_UNSAT_OTHER_PREV_YEAR);
      USE(WG_USE_REDUCED_RETURN) -= USE(WG_USE_REDUCED_RETURN_PREV_YEAR);
      if ((USE(WG_USE_UNSAT_IRRIG) + USE(WG_USE_UNSAT_OTHER)) < 0) {
        USE(WG_USE_UNSAT_IRRIG) = 0;
        USE(WG_USE_UNSAT_OTHER) = 0;
        USE(WG_USE_REDUCED_RETURN) = 0;
      } else if (USE(WG_USE_UNSAT_OTHER) < 0) {
        USE(WG_USE_UNSAT_IRRIG) += USE(WG_USE_UNSAT_OTHER);
        USE(WG_USE_UNSAT_OTHER) = 0;
      } else if (USE(WG_USE_UNSAT_IRRIG) < 0) {
        USE(WG_USE_UNSAT_OTHER) += USE(WG_USE_UNSAT_IRRIG);
        USE(WG_USE_UNSAT_IRRIG) = 0;
        USE(WG_USE_REDUCED_RETURN) = 0;
      }
      if (USE(WG_USE_REDUCED_RETURN) > 0) {
        USE(WG_USE_REDUCED_RETURN) = 0;
        USE(WG_USE_UNSAT_IRRIG) = 0;
      }
    }
  }
}

void wgo_get_water_use_state(const wgo_model *m, double *out) {
  if (m->use) memcpy(out, m->use, (size_t)m->n * WG_USE_NFIELDS * sizeof(double));
}

void wgo_set_water_use_state(wgo_model *m, const double *in) {
  if (m->use) memcpy(m->use, in, (size_t)m->n * WG_USE_NFIELDS * sizeof(double));
}

/* ------------------------------------------------------------------ */
/* ensemble forcing perturbation (no reference counterpart)             */
/* ------------------------------------------------------------------ */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), restated from the paper:
 * round multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments 0x9E3779B9 /
 * 0xBB67AE85, ten rounds. */
static void philox4x32_10(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t c0 = hi1 ^ ctr[1] ^ k0, c2 = hi0 ^ ctr[3] ^ k1;
    ctr[0] = c0;
    ctr[1] = lo1;
    ctr[2] = c2;
    ctr[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

/* Two independent standard normal deviates for (cell, day) of one member by
 * Box-Muller from one Philox block: u = (c + 1/2) 2^-32, R = sqrt(-2 ln u1),
 * (R cos 2 pi u2, R sin 2 pi u2), with the log, sin and cos shared with the
 * kernels (wg_libm.h) and the correctly rounded sqrt. */
static void normal_pair(uint64_t seed, uint32_t cell, uint32_t day, double *z0, double *z1) {
  uint32_t c[4] = {cell, day, 0x424f584du, 0u};
  double u1, u2, r, sn, cs;
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  u1 = ((double)c[0] + 0.5) * 2.3283064365386963e-10;
  u2 = ((double)c[1] + 0.5) * 2.3283064365386963e-10;
  r = sqrt(-2.0 * wg_log(u1));
  wg_sincos2pi(u2, &sn, &cs);
  *z0 = r * cs;
  *z1 = r * sn;
}

void wgo_normal_pair(uint64_t seed, uint32_t cell, uint32_t day, double *out2) {
  normal_pair(seed, cell, day, out2, out2 + 1);
}

void wgo_perturb_forcing(uint64_t seed, int32_t day_index, int32_t n, double prec_sigma,
                         double temp_sigma, const float *temp_in, const float *prec_in,
                         float *temp_out, float *prec_out) {
  for (int32_t i = 0; i < n; i++) {
    double zp, zt;
    normal_pair(seed, (uint32_t)i, (uint32_t)day_index, &zp, &zt);
    prec_out[i] = (float)((double)prec_in[i] * wg_exp(prec_sigma * zp));
    temp_out[i] = (float)((double)temp_in[i] + temp_sigma * zt);
  }
}

/* One Philox4x32-10 block, exported for the known-answer test. */
void wgo_philox(uint32_t ctr[4], const uint32_t key[2]) { philox4x32_10(ctr, key[0], key[1]); }

/* The math functions in use, exported for the tests. */
double wgo_exp(double x) { return EXP(x); }
double wgo_pow(double x, double y) { return POW(x, y); }
double wgo_log(double x) { return wg_log(x); }
void wgo_sincos2pi(double u, double *s, double *c) { wg_sincos2pi(u, s, c); }
