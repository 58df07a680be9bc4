/*
 * watergap2_amd.h -- C ABI of the MI355X implementation of the WaterGAP 2 daily
 * grid-cell water balance and lateral routing.
 *
 * Every entry point has C linkage, takes plain pointers and sizes, returns 0 on
 * success or a negative WG_E_* code, and never lets an exception cross the
 * boundary.  wg_last_error() gives the text of the last failure.
 *
 * Cell arrays are in the reference's cell order n = 0..N-1 (the order of the
 * reference's Grid<> containers, grid.h:58-136).  Per-cell channel arrays are
 * cell-major, channel-contiguous like VariableChannelGrid (grid.h:471-474):
 * elevation[n*101 + ch].  The library permutes to its own device layout
 * internally; callers never see it.
 *
 * Reference interfaces replaced (file:line in HydrologyFrankfurt/WaterGAP2
 * source/):
 *   wg_rout_order            rout_prepare.cpp:834-886  rout_order()
 *   wg_inflow_outflow        rout_prepare.cpp:338-439, 619-670  inflow_cells(), outflow_cell()
 *   wg_create                routing.cpp:131-742 routingClass::init, daily.cpp:1544-1682
 *                            dailyWaterBalanceClass::init, lai.cpp:40-108 laiClass::init,
 *                            routing.cpp:745-765 initFractionStatus, routing.cpp:5621-5720
 *                            initLake/WetlDepthActive + setLakeWetlToMaximum,
 *                            routing.cpp:979-1495 annualInit (first year)
 *   wg_get_state/wg_set_state daily.cpp:1896-1924 setStorages, routing.cpp:851-882 setStorages
 *   wg_load_forcing          climateYear.cpp:57-79 (daily .365 forcing grids)
 *   wg_step_days             integrateWGHM.cpp:755-798: per day
 *                            dailyWaterBalanceClass::calcNewDay (daily.cpp:94-1279) for all n,
 *                            routingClass::routing (routing.cpp:1629-5244),
 *                            routingClass::updateLandAreaFrac (routing.cpp:5343-5352)
 *   wg_get_station_discharge routing.cpp:4232-4238 dailyRiverDischarge, 7232-7239 getAnnualDischarge
 *   wg_destroy               routingClass/dailyWaterBalanceClass destructors
 */
#ifndef WATERGAP2_AMD_H
#define WATERGAP2_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WG_ABI_VERSION 2

/* error codes */
#define WG_OK 0
#define WG_E_ARG (-1)       /* invalid argument (null pointer, size out of range) */
#define WG_E_CYCLE (-2)     /* flow-direction graph is cyclic or dangling (rout_prepare.cpp:876-881) */
#define WG_E_HIP (-3)       /* HIP runtime error */
#define WG_E_NOMEM (-4)     /* allocation failed */
#define WG_E_INPUT (-5)     /* invalid input value (e.g. arid index not 0/1, daily.cpp:345-347) */
#define WG_E_STATE (-6)     /* call out of sequence */

#define WG_NSUB 100          /* elevation sub-cells per cell (daily.cpp:920) */
#define WG_NELEV 101         /* channels of G_Elevation / G_SnowInElevation (daily.h:49,122) */
#define WG_NLCT 18           /* land cover classes (def.h:20) */

/* Calibration parameters, in the order of eCalibParam (calib_param.h:72-98). */
enum wg_param {
  WG_P_GAMRUN_C = 0, WG_P_CFA, WG_P_CFS, WG_M_ROOT_D, WG_M_RIVRGH_C, WG_P_LAK_D, WG_P_WET_D,
  WG_P_SWOUTF_C, WG_M_EVAREDEX, WG_M_NETRAD, WG_P_PTC_HUM, WG_P_PTC_ARI, WG_P_PET_MXDY, WG_P_MCWH,
  WG_M_LAI, WG_P_T_SNOWFZ, WG_P_T_SNOWMT, WG_M_DEGDAY_F, WG_P_T_GRADNT, WG_M_GW_F, WG_M_RG_MAX,
  WG_P_PCRITGWA, WG_P_GWOUTF_C, WG_M_NETABSSW, WG_M_NETABSGW, WG_M_PREC, WG_NPARAM
};

/* Land cover tables: LCT_22.DAT (daily.cpp:1925-1959) and LAI_22.DAT (lai.cpp:110-149). */
typedef struct wg_lct_tables {
  float rooting_depth[WG_NLCT];
  double albedo[WG_NLCT];
  double albedo_snow[WG_NLCT];
  double ddf[WG_NLCT];
  double emissivity[WG_NLCT];
  float lai[WG_NLCT];
  float decid_frac[WG_NLCT];
  float evergreen_red[WG_NLCT];
  int16_t initial_days[WG_NLCT];
  double kc_min[WG_NLCT];
  double kc_max[WG_NLCT];
} wg_lct_tables;

/*
 * Static per-cell inputs with the reference's own element types.  All arrays
 * have N entries unless stated.  Raw inputs, as the reference reads them from
 * its .UNF files; wg_create derives the routing geometry from them the way
 * routingClass::init does (routing.cpp:476-507).
 */
typedef struct wg_static_fields {
  /* geometry: geo.cpp:7-81 */
  const int16_t *row;          /* GR: 1-based row, row 1 = 89.75 N */
  const double *cell_area;     /* GAREA by row, km2 (geo.h:39-41) */
  const double *contfreq;      /* GCONTFREQ, % continental area */
  /* vertical balance: daily.cpp, lai.cpp, gw_frac.cpp, s_max.cpp, land.h */
  const int16_t *aindex;       /* G_aindex: 1 arid, 0 humid */
  const int8_t *ldd;           /* G_LDD: 1-9 D8 (keypad), 5 sink/outlet, -1 inland sink */
  const int8_t *land_cover;    /* 1..18 */
  const float *built_up;       /* land.G_built_up fraction */
  const float *smax;           /* maxSoilWaterCap.G_Smax, mm */
  const float *gw_factor;      /* GW.getgwFactor */
  const int16_t *rgmax;        /* GW.getRgmax, mm/d * 100 */
  const int8_t *texture;       /* GW.G_texture */
  const int16_t *elevation;    /* [N][101] G_Elevation, m; channel 0 = cell mean */
  const double *params;        /* [WG_NPARAM][N] calibration parameters */
  /* routing: routing.cpp:131-507 */
  const double *glo_lake;      /* G_GLOLAK % */
  const double *loc_lake;      /* G_LOCLAK % (local reservoirs already added) */
  const double *glo_wetland;   /* G_GLOWET % */
  const double *loc_wetland;   /* G_LOCWET % */
  const double *lake_area;     /* G_LAKAREA km2 */
  const double *reservoir_area;/* G_RESAREA km2 of operating reservoirs */
  const double *glo_res;       /* G_RES % */
  const double *stor_cap;      /* G_STORAGE_CAPACITY km3 */
  const double *mean_outflow;  /* G_MEAN_OUTFLOW converted to m3/s (routing.cpp:424-425) */
  const int8_t *res_type;      /* 1 irrigation, 2 other */
  const int8_t *start_month;   /* 1..12, start of operational year */
  const float *river_length;   /* G_RIVER_LENGTH km, before the contfreq scaling */
  const float *river_slope;    /* G_RIVERSLOPE m/m */
  const float *roughness;      /* G_ROUGHNESS */
  const float *bankfull_flow;  /* G_BANKFULL m3/s */
  const int32_t *downstream;   /* G_OUTFLC: 1-based downstream cell, 0 = none */
  const int32_t *rout_order;   /* G_ROUT_ORDER: 1-based position of each cell in the routing loop */
  const wg_lct_tables *lct;
  /* mean demand of irrigation reservoirs (res_type 1), routing.cpp:344-403; both may be NULL */
  const double *mean_nus;      /* G_NUs_<first>_<last>.UNF0: mean annual net surface-water use, m3/yr
                                  (NULL: no demand, so type-1 reservoirs release like type 2) */
  const double *alloc_coeff;   /* [N][5] G_ALLOC_COEFF.5.UNF0 (rout_prepare.cpp:895-944, wg_alloc_coeff);
                                  NULL: 1 for every cell */
} wg_static_fields;

/* Option profile.  Only the values below are implemented; wg_create rejects others. */
typedef struct wg_options_profile {
  int32_t time_series;          /* 1: .365 forcing */
  int32_t cloud;                /* 1: downward LW + emissivity */
  int32_t intercept;            /* 1 */
  int32_t calc_albedo;          /* 0 */
  int32_t pet_opt;              /* 0: Priestley-Taylor */
  int32_t use_kc;               /* 0 */
  int32_t river_velo_opt;       /* 2: storage-based Manning velocity */
  int32_t subtract_use;         /* 0; 2 or 3 (water use, wg_set_water_use) with aggr_nus_glolakres_opt 0 */
  int32_t res_opt;              /* 1 */
  int32_t statcorr_opt;         /* 0 */
  int32_t arid_area_opt;        /* 1 */
  int32_t fractional_routing_opt; /* 1 */
  int32_t river_evapo_opt;      /* 1 */
  int32_t aggr_nus_glolakres_opt; /* 0 or 1; identical results without water use */
} wg_options_profile;

/* Fills the parity option profile. */
void wg_default_options(wg_options_profile *opt);

/*
 * Prognostic state of one member, reference cell order.  Storages of the
 * vertical balance in mm, routing storages in km3 (as the reference keeps them).
 * Arrays of N entries, snow_elev of N*101.
 */
typedef struct wg_state {
  double *canopy;        /* G_canopyWaterContent */
  double *soil;          /* G_soilWaterContent */
  double *snow;          /* G_snow */
  double *snow_elev;     /* [N][101] G_SnowInElevation (channel 0 unused) */
  double *storage_transfer; /* G_dailyStorageTransfer */
  int32_t *lai_days;     /* G_days_since_start */
  int32_t *lai_status;   /* G_GrowingStatus */
  double *lai_precsum;   /* G_PrecSum */
  double *groundwater, *loc_lake, *loc_wetl, *glo_lake, *glo_wetl, *reservoir, *river;
  double *red_loc_lake;  /* G_locLakeAreaReductionFactor */
  double *red_loc_wetl;  /* G_locWetlAreaReductionFactor */
  double *red_glo_lake;  /* G_gloLakeEvapoReductionFactor */
  double *red_glo_wetl;  /* G_gloWetlAreaReductionFactor */
  double *red_res;       /* G_gloResEvapoReductionFactor */
  double *red_river;     /* G_riverAreaReductionFactor */
  double *river_frac_next; /* G_riverAreaFracNextTimestep_Frac */
  double *fswb_frac_next;  /* G_fswbLandAreaFracNextTimestep */
  double *laf;           /* G_landAreaFrac */
  double *laf_prev;      /* G_landAreaFracPrevTimestep */
  double *k_release;     /* K_release */
  int32_t started;       /* statusStarted_landAreaFracNextTimestep (1 after the first routed day) */
} wg_state;

/* Per-cell daily diagnostics of the last simulated day (optional). */
typedef struct wg_day_outputs {
  double *discharge;     /* transportedVolume km3/d (routing.cpp:3955-3958) */
  double *surface_runoff;/* G_dailyLocalSurfaceRunoff mm */
  double *gw_recharge;   /* G_dailyGwRecharge mm */
  double *open_water_pet;/* G_openWaterPET mm */
  double *cell_aet_uncorr; /* dailyCellLandAET_uncorr mm */
  double *prec;          /* precipitation after M_PREC, mm */
} wg_day_outputs;

typedef struct wg_handle wg_handle;

/* ---- routing schedule (host) ---- */

/* Inflow cells (9 channels, 1-based, 0 = none) and outflow cell (1-based, 0 =
 * none) from the LDD on a lattice.  cell_at[(row-1)*720 + (col-1)] gives the
 * 1-based cell at each 0.5-degree position (0 = no land), rows 1..360, columns
 * 1..720 with wrap in longitude.  Replaces rout_prepare.cpp:338-439, 619-670. */
int wg_inflow_outflow(int32_t n_cells, const int8_t *ldd, const int16_t *row, const int16_t *col,
                      const int32_t *cell_at, int32_t *inflow9, int32_t *outflow);

/* Routing order (1-based) and Kahn sweep (level, 0-based) from the 9-channel
 * inflow cells and the outflow cell.  Returns WG_E_CYCLE when some cell cannot
 * be ordered.  Replaces rout_prepare.cpp:834-886; level may be NULL. */
int wg_rout_order(int32_t n_cells, const int32_t *inflow9, const int32_t *outflow,
                  int32_t *rout_order, int32_t *level);

/* ---- WaterGAP input files (host) ----
 * .UNF grids are raw big-endian arrays without a header (grid_io_adapters.h:43-129);
 * the digit after .UNF is the element size: 0 float32, 1 int8, 2 int16, 4 int32.
 * The reference keeps double grids in .UNF0 files as float32. */
#define WG_UNF_F32 0
#define WG_UNF_I8 1
#define WG_UNF_I16 2
#define WG_UNF_I32 4

/* Element type of a file from its name (WG_UNF_*), or WG_E_ARG. */
int wg_unf_type(const char *path);
/* Reads count elements of the given type into out (host byte order); WG_E_INPUT
 * when the file is missing or shorter.  Replaces GridUNFAdapter::readGrid for
 * grids of the file's own element type. */
int wg_unf_read(const char *path, int32_t type, int64_t count, void *out);
int wg_unf_write(const char *path, int32_t type, int64_t count, const void *in);
/* A double grid read from / written to a .UNF0 file (float32 on disk), as
 * GridUNFAdapter does for Grid<double> (grid_io_adapters.h:60-75, 103-117). */
int wg_unf_read_double(const char *path, int64_t count, double *out);
int wg_unf_write_double(const char *path, int64_t count, const double *in);
/* The year's daily forcing from climate_dir/G_TEMP_H08_int_<year>.365.UNF0,
 * G_GPCC_H08day_V20110128_<year>.365.UNF0, G_SSRD_H08_int_<year>.365.UNF0 and
 * G_SLRD_H08_int_<year>.365.UNF0 ([N][365] float each) into the layout of
 * wg_load_forcing, forcing[365][4][N].  Replaces
 * climateYearClass::read_climate_data_daily_per_year (climateYear.cpp:38-54). */
int wg_read_climate_year(const char *climate_dir, int32_t year, int32_t n_cells, float *forcing);
/* LCT_22.DAT and LAI_22.DAT of input_dir ('#' comment lines, then one row per
 * class 1..18).  Replaces dailyWaterBalanceClass::readLCTdata (daily.cpp:1925-1959)
 * and laiClass::readLAIdata (lai.cpp:110-149). */
int wg_read_lct_tables(const char *input_dir, wg_lct_tables *out);
/* Cell area from GAREA.UNF0 by row (geo.h:33-35) and the continental share raised
 * to the sum of the surface-water fractions G_LOCLAK, G_GLOLAK, G_LOCRES,
 * G_RES_FRAC (any may be NULL).  Replaces geo.cpp:7-44. */
int wg_cell_geometry(int32_t n_cells, const double *area_by_row, const int16_t *row, const double *contfreq_in,
                     const double *loc_lake, const double *glo_lake, const double *loc_res, const double *res_frac,
                     double *cell_area, double *contfreq);
/* Smax = G_TAWC x M_ROOT_D x rooting depth of the land-cover class (-9999 where
 * TAWC < 0).  Replaces soilWatCapClass::createMaxSoilWaterCapacityGrid
 * (s_max.cpp:43-74). */
int wg_soil_capacity(int32_t n_cells, const float *tawc, const int8_t *land_cover, const double *m_root_d,
                     const float *rooting_depth_lct, float *smax);
/* Groundwater factor fg and Rgmax (mm/d x 100) from G_SLOPE_CLASS.UNF1,
 * G_TEXTURE.UNF1, G_PERMAGLAC.UNF1, G_AQ_FACTOR.UNF1 and G_GW_FACTOR_CORR.UNF0 (may
 * be NULL) with the M_GW_F and M_RG_MAX multipliers; time_series 0 selects the
 * daily-WFD Rgmax table.  Replaces groundwaterFactorClass::createGrids
 * (gw_frac.cpp:31-258). */
int wg_gw_factor(int32_t n_cells, const int8_t *slope_class, const int8_t *texture, const int8_t *permaglac,
                 const int8_t *aquifer_factor, const float *gw_factor_corr, const double *m_gw_f,
                 const double *m_rg_max, int32_t time_series, float *gw_factor, int16_t *rgmax);

/* ---- restart files (host, text variants) ---- */
#define WG_NSTORE 10        /* canopy, snow, soil, local lake, local wetland, global lake, global
                               wetland, reservoir, river, groundwater (wghmStateFile.h) */
#define WG_NADDITIONAL 53   /* columns of the additional out- and input file */

/* WGHM water storage states, mm over each cell's continental area [N][WG_NSTORE]
 * (TWS, their sum, is written first and skipped on reading).  Replace
 * WghmStateFile::saveMean / load for .txt names (wghmStateFile.cpp:156-229, 357-413). */
int wg_write_wghm_state_txt(const char *path, int32_t n_cells, const double *storages_mm);
int wg_read_wghm_state_txt(const char *path, int32_t n_cells, double *storages_mm);
/* Snow of the 101 elevation channels, mm over the continental area [N][101].  Replace
 * SnowInElevationFile::save / load for .txt names (snowInElevationFile.cpp:80-210). */
int wg_write_snow_elevation_txt(const char *path, int32_t n_cells, const double *snow);
int wg_read_snow_elevation_txt(const char *path, int32_t n_cells, double *snow);
/* The 53 columns of the additional out- and input file [N][53].  Replace
 * AdditionalOutputInputFile::save / load for .txt names
 * (additionalOutputInputFile.cpp:186-400). */
int wg_write_additional_txt(const char *path, int32_t n_cells, const double *cols);
int wg_read_additional_txt(const char *path, int32_t n_cells, double *cols);
/* A member's state to the three files' values and back: storages in mm over the
 * continental area (routing.cpp:4998-5020, integrateWGHM.cpp:830-846;
 * daily.cpp:1896-1924, routing.cpp:851-882), the LAI state, release coefficient,
 * land-area fractions and reduction factors in the additional columns (lai.cpp:55-64,
 * 165-173; routing.cpp:388-393, 3080, 5209-5233).  Any output may be NULL; NULL
 * state arrays are skipped.  The river's reduction factor and area fraction and the
 * storage transfer are not in the reference's files and are left as they are. */
int wg_state_to_files(int32_t n_cells, const wg_static_fields *st, const wg_state *state, double *storages_mm,
                      double *snow_mm, double *additional);
int wg_state_from_files(int32_t n_cells, const wg_static_fields *st, const double *storages_mm,
                        const double *snow_mm, const double *additional, wg_state *state);

/* ---- routing preprocessing (host): prepare_routing_files, rout_prepare.cpp:61-336 ----
 * Cell arrays in the reference's cell order; cell numbers in inflow9 / outflow are
 * 1-based with 0 = none, as in G_INFLC.9.UNF4 / G_OUTFLC.UNF4. */

/* Arc/Info flow codes (G_FLOWDIR.UNF2: 1 E, 2 SE, 4 S, 8 SW, 16 W, 32 NW, 64 N,
 * 128 NE, 0 sink, -1 inland sink) to WaterGAP LDD codes; any other value is a sink.
 * Replaces rout_prepare.cpp:68-112. */
int wg_ldd_from_flowdir(int32_t n_cells, const int16_t *flowdir, int8_t *ldd);
/* Cells whose flow direction points at a position without land become sinks
 * (LDD 5).  Replaces rout_prepare.cpp:287-336 (replace_flow_to_ocean, which the
 * reference keeps but prepare_routing_files does not call). */
int wg_replace_flow_to_ocean(int32_t n_cells, const int16_t *row, const int16_t *col, const int32_t *cell_at,
                             int8_t *ldd);
/* Number of cells draining through each cell, itself included (G_FLOW_ACC.UNF2,
 * 16-bit); 0 for cells on or below a loop.  Replaces rout_prepare.cpp:441-494. */
int wg_flow_accumulation(int32_t n_cells, const int32_t *inflow9, int16_t *flow_acc);
/* Basin number of every cell (endpoints LDD 5 / -1 numbered 1.. in cell order,
 * G_BASINS.UNF2) and the number of cells to the outlet (G_CELLS_TO_OUTLET.UNF2,
 * may be NULL); LDD codes 0 and -99 become 5, as in the reference.  Replaces
 * rout_prepare.cpp:497-571. */
int wg_derive_basins(int32_t n_cells, int8_t *ldd, const int32_t *inflow9, uint16_t *basin, uint16_t *dist,
                     int32_t *n_basins);
/* Renumbers the basins of more than one cell 1..n_multi and sets the one-cell
 * basins to 0 (G_BASINS_2.UNF2).  Replaces rout_prepare.cpp:573-604. */
int wg_reindex_basins(int32_t n_cells, int32_t n_basins, uint16_t *basin, int32_t *n_multi);
/* The 8 neighbour cells of each cell (E, NE, N, NW, W, SW, S, SE; 0 = none,
 * G_NEIGHBOUR_CELLS.8.UNF4).  Replaces rout_prepare.cpp:672-692. */
int wg_neighbour_cells(int32_t n_cells, const int16_t *row, const int16_t *col, const int32_t *cell_at,
                       int32_t *neighbour8);
/* Distances (km) from a cell of row r (1-based) to its neighbour in LDD direction
 * k (1..9): dist[(k-1)*360 + (r-1)] (GCELLDIST.9.UNF0).  Replaces
 * rout_prepare.cpp:694-752. */
int wg_cell_distances(float *dist);
/* River slope (m/m, floor 1e-5) along the flow direction from GALTMOD.UNF0 and the
 * meandering ratio (G_RIVERSLOPE.UNF0).  Replaces rout_prepare.cpp:754-788. */
int wg_river_slope(int32_t n_cells, const float *meander, const float *dist, const float *altitude,
                   const int32_t *outflow, const int8_t *ldd, const int16_t *row, float *slope);
/* River length (km) with the meandering ratio (G_RIVER_LENGTH.UNF0).  Replaces
 * rout_prepare.cpp:790-832. */
int wg_river_length(int32_t n_cells, const float *meander, const int16_t *row, const int16_t *col, const float *dist,
                    const int32_t *outflow, float *length);
/* Allocation coefficients [N][5] of the up to 5 cells downstream of each
 * reservoir (G_ALLOC_COEFF.5.UNF0; the reference stores them as float); 1 where no
 * reservoir allocates.  mean_outflow as read from G_MEAN_OUTFLOW.UNF0.  Replaces
 * rout_prepare.cpp:895-944. */
int wg_alloc_coeff(int32_t n_cells, const int32_t *outflow, const double *reservoir_area, const double *mean_outflow,
                   double *alloc_coeff);
/* Start month (1..12) of the reservoirs' operational year from the monthly and
 * annual mean outflow (G_START_MONTH.UNF1).  Replaces rout_prepare.cpp:950-1019. */
int wg_reservoir_start_month(int32_t n_cells, const double *mean_outflow_monthly, const double *mean_outflow,
                             int8_t *start_month);

/* Inputs and outputs of the whole preprocessing step (file names of the reference). */
typedef struct wg_routing_inputs {
  int32_t n_cells;
  const int16_t *flowdir;       /* G_FLOWDIR.UNF2 (Arc/Info codes), or NULL to use ldd */
  const int8_t *ldd;            /* G_LDD.UNF1 when flowdir is NULL */
  const int16_t *row, *col;     /* GR.UNF2, GC.UNF2 */
  const int32_t *cell_at;       /* GCRC.UNF4 as [row-1][col-1] of 360 x 720, 1-based cell, 0 = ocean */
  const float *altitude;        /* GALTMOD.UNF0 */
  const float *meander;         /* G_MEANDERING_RATIO.UNF0 */
  const double *reservoir_area; /* G_RESAREA.UNF0 (NULL: no reservoir outputs) */
  const double *mean_outflow;   /* G_MEAN_OUTFLOW.UNF0 */
  const double *mean_outflow_monthly; /* G_MEAN_OUTFLOW.12.UNF0 [N][12] */
} wg_routing_inputs;

typedef struct wg_routing_files {
  int8_t *ldd;                  /* flow directions after the loop and no-data corrections */
  int32_t *inflow9;             /* G_INFLC.9.UNF4 [N][9] */
  int32_t *outflow;             /* G_OUTFLC.UNF4 */
  int16_t *flow_acc;            /* G_FLOW_ACC.UNF2 */
  uint16_t *basins;             /* G_BASINS_2.UNF2 */
  uint16_t *cells_to_outlet;    /* G_CELLS_TO_OUTLET.UNF2 (may be NULL) */
  float *cell_dist;             /* GCELLDIST.9.UNF0 [9][360] (may be NULL) */
  float *river_slope;           /* G_RIVERSLOPE.UNF0 (may be NULL) */
  float *river_length;          /* G_RIVER_LENGTH.UNF0 (may be NULL) */
  int32_t *rout_order;          /* G_ROUT_ORDER.UNF4 */
  double *alloc_coeff;          /* G_ALLOC_COEFF.5.UNF0 [N][5] (may be NULL) */
  int8_t *start_month;          /* G_START_MONTH.UNF1 (may be NULL) */
  int32_t n_basins;             /* basins of more than one cell */
} wg_routing_files;

/* prepare_routing_files (rout_prepare.cpp:61-336): flow directions, inflow cells,
 * flow accumulation with the loop correction, basins, outflow cells, distances,
 * slope, length, routing order and, with reservoir inputs, the allocation
 * coefficients and start months. */
int wg_prepare_routing(const wg_routing_inputs *in, wg_routing_files *out);

/* ---- model ---- */

/* Creates a model of n_members members sharing the static fields.  Reservoirs
 * with reservoir_area > 0 are in operation from the first day (resYearOpt = 1,
 * routing.cpp:5647-5720).  1 <= n_cells < 2^27; the cells that drain to one outlet, packed
 * into routing groups, must fit the LDS limit of 18,204 cells per group, 1 <= n_members <= 65535. */
int wg_create(int32_t n_cells, const wg_static_fields *st, const wg_options_profile *opt,
              int32_t n_members, wg_handle **out);
int wg_destroy(wg_handle *h);
const char *wg_last_error(const wg_handle *h);

/* State of one member (member -1 in wg_set_state: every member).  NULL array
 * pointers in the struct are skipped. */
int wg_get_state(wg_handle *h, int32_t member, wg_state *out);
int wg_set_state(wg_handle *h, int32_t member, const wg_state *in);

/* Per-member values of the runoff coefficient gamma, the cell correction factor
 * CFA and the soil capacity Smax (ensemble members and calibration candidates).
 * Arrays are [n_members][N] in reference cell order; NULL keeps the current values. */
int wg_set_member_params(wg_handle *h, const double *gamma, const double *cfa, const float *smax);

/* Per-member forcing perturbation of the ensemble generator: seed per member,
 * precipitation lognormal sigma and temperature normal sigma.  sigma 0 disables. */
int wg_set_forcing_perturbation(wg_handle *h, const uint64_t *member_seed, double prec_sigma,
                                double temp_sigma);

/* Copies n_days of daily forcing [n_days][4][N] (temperature degC, precipitation
 * mm/d, shortwave W/m2, longwave W/m2; reference cell order) into device memory.
 * WG_E_INPUT, with nothing changed, when a value is not finite or a precipitation
 * is negative. */
int wg_load_forcing(wg_handle *h, const float *forcing, int32_t n_days);

/* routingClass::annualInit for a year after the first (routing.cpp:979-1495), at
 * a year boundary between step calls.  st holds the statics of wg_create with this
 * year's reservoir grids as Parts 1-2 of annualInit leave them (routing.cpp:1015-1100:
 * G_RES of the year to use, and the area and capacity of the reservoirs in operation;
 * wg_reservoir_year computes the latter); every other field must be the one given to
 * wg_create.  With apply_changes (the reference's condition start month == 1 or a
 * year after the first, routing.cpp:1313), a cell whose reservoir fraction grew
 * against the previous call (or wg_create) loses that fraction of land, and the
 * canopy, soil and snow water of the lost land moves to the reservoir storage of
 * outflow_cell[i] (1-based, G_OUTFLOW_CELL_ASSIGNMENT.UNF4), every member
 * (routing.cpp:1300-1421).  WG_E_INPUT, with nothing changed, when such a cell's land
 * fraction does not fall against the day before (routing.cpp:1354-1357, a branch this
 * state does not represent) or a reservoir has an unknown type. */
int wg_annual_init(wg_handle *h, const wg_static_fields *st, const int32_t *outflow_cell, int32_t apply_changes);
/* Parts 1-2 of annualInit (routing.cpp:1015-1100) for the year to use (resYearCurrentToUse):
 * regulated lakes (reg_lake_status 1) and reservoirs with start_year <= year_to_use
 * operate with their full area and capacity, the rest have none. */
int wg_reservoir_year(int32_t n_cells, int32_t year_to_use, const double *reservoir_area_full,
                      const double *stor_cap_full, const int32_t *start_year, const int8_t *reg_lake_status,
                      double *reservoir_area, double *stor_cap);

/* Station (outlet) cells whose daily discharge is accumulated (0-based cells). */
int wg_set_stations(wg_handle *h, int32_t n_stations, const int32_t *cells);

/* Runs n_days days: forcing day forcing_day0+i, julian day doy (1..365)
 * starting at first_doy and wrapping after 365. */
int wg_step_days(wg_handle *h, int32_t forcing_day0, int32_t first_doy, int32_t n_days);

/* Station discharge summed over the days since the last reset [n_members][n_stations]
 * (km3), and, when daily is not NULL, the daily series kept from the last
 * wg_step_days / wg_step_days_profiled call: [n_members][max_days][n_stations], of
 * which the first min(max_days, kept days) rows are written.  wg_step_days runs
 * blocks of up to 365 days and keeps the series of its last block;
 * wg_station_daily_window tells which days those are. */
int wg_get_station_discharge(wg_handle *h, double *sum, double *daily, int32_t max_days);
/* The days of the kept daily series: first_day counted from the first day of the
 * last step call (0 for calls of up to 365 days), n_days of them.  No counterpart
 * in the reference, whose daily series is a whole year (routing.cpp:4232-4238). */
int wg_station_daily_window(wg_handle *h, int32_t *first_day, int32_t *n_days);
int wg_reset_station_discharge(wg_handle *h);

/* G_potCellRunoff (routing.h:286), the per-cell sum of the daily potential cell
 * runoff in km3 (routing.cpp:1962-1970 (semi-)arid cells through 3372-3379, 2045-2053
 * humid cells, 2180-2187 inland sinks): the cell's local runoff terms plus the
 * open-water balance (daily.cpp:773) over its lake, reservoir and wetland area, the
 * quantity the cell correction factors are computed from.  An output the day loop
 * keeps only when asked: enable 1 allocates it (or, when on, restarts the sums at
 * 0), 0 frees it.  wg_annual_init restarts the sums (routing.cpp:1009). */
int wg_set_pot_cell_runoff(wg_handle *h, int32_t enable);
/* The sums since the last restart for one member, pot[N] in reference cell order. */
int wg_get_pot_cell_runoff(wg_handle *h, int32_t member, double *pot);

/*
 * Water use (subtract_use 2 or 3 with use_alloc = 0 and aggrNUsGloLakResOpt = 0):
 * net abstractions from surface water (NUs) and groundwater (NUg) taken from the
 * cell's own storages, in the reference's order: the groundwater abstraction,
 * adapted by updateNetAbstractionGW (routing.cpp:5503-5572), in the groundwater
 * step (1929-2137, 3325-3330); the surface-water use from the global lake and
 * reservoir (2678-3171), the river (3466-3512) and last the local lake
 * (3590-3622); the unsatisfied use kept for the next day (3898-3919) and the
 * year-end processing (annualWaterUsePostProcessing, 5246-5306) after day 365.
 * The irrigation reservoirs' release takes the month's use of the cell and its
 * downstream cells (2970-2996).
 *
 * Per-cell state of one member, [WG_USE_NFIELDS][N] in reference cell order
 * (km3; routing.h:204-318).
 */
enum {
  WG_USE_TOTAL_UNSAT = 0,       /* G_totalUnsatisfiedUse */
  WG_USE_PREV_TOTAL_UNSAT,      /* G_PrevTotalUnstatisfiedUse */
  WG_USE_DAILY_REMAINING,       /* G_dailyRemainingUse */
  WG_USE_UNSAT_IRRIG,           /* G_unsatisfiedNAsFromIrrig */
  WG_USE_UNSAT_OTHER,           /* G_unsatisfiedNAsFromOtherSectors */
  WG_USE_REDUCED_RETURN,        /* G_reducedReturnFlow */
  WG_USE_WITHDRAWAL_IRRIG,      /* G_withdrawalIrrigFromSwb (km3/day) */
  WG_USE_CONSUMPTIVE_IRRIG,     /* G_consumptiveUseIrrigFromSwb (km3/day) */
  WG_USE_UNSAT_PREV_YEAR,       /* G_UnsatisfiedUsePrevYear */
  WG_USE_UNSAT_IRRIG_PREV_YEAR, /* G_unsatisfiedNAsFromIrrigPrevYear */
  WG_USE_UNSAT_OTHER_PREV_YEAR, /* G_unsatisfiedNAsFromOtherSectorsPrevYear */
  WG_USE_REDUCED_RETURN_PREV_YEAR, /* G_reducedReturnFlowPrevYear */
  WG_USE_NFIELDS
};

/* Turns water use on for a handle made with subtract_use 2 or 3 (WG_E_STATE
 * otherwise).  Tables of n_years model years, [n_years][N][12] in reference cell
 * order as the .12.UNF0 files hold them, in m3 per month: net use from surface
 * water (G_NETUSE_SW_m3_<year> or G_ACTUAL_NAS_<year>) and from groundwater,
 * irrigation withdrawal and consumptive use from surface water
 * (G_IRRIG_WITHDRAWAL_USE_SW_m3_<year>, G_IRRIG_CONS_USE_SW_m3_<year>).  Year y
 * covers the forcing days year0_day + 365 y .. + 364; wg_step_days refuses a day
 * outside the tables.  frac_return_gw [N] is G_FRACTRETURNGW_IRRIG.  delayed is
 * delayedUseSatisfaction (0 or 1).  Daily values as dailyNUInit
 * (routing.cpp:909-972): times M_NETABSSW / M_NETABSGW, over 1e9 x days in the
 * month.  st: the statics of wg_create or of the last wg_annual_init (the
 * irrigation reservoirs' release walks their downstream cells); after a
 * wg_annual_init the tables must be set again (wg_step_days: WG_E_STATE).  The
 * state starts at 0; wg_set_water_use_state sets it. */
int wg_set_water_use(wg_handle *h, const wg_static_fields *st, int32_t n_years, int32_t year0_day, const double *nus_m3,
                     const double *nug_m3, const double *withdrawal_irrig_m3, const double *consumptive_irrig_m3,
                     const double *frac_return_gw, int32_t delayed);
int wg_get_water_use_state(wg_handle *h, int32_t member, double *state);
int wg_set_water_use_state(wg_handle *h, int32_t member, const double *state);

/* Diagnostics of the last simulated day for one member. */
int wg_get_day_outputs(wg_handle *h, int32_t member, wg_day_outputs *out);

/* Synchronises the handle's stream. */
int wg_synchronize(wg_handle *h);

/* Device id and stream (hipStream_t as void*) the handle launches on. */
void *wg_stream(wg_handle *h);

/* As wg_step_days, launched eagerly with a HIP event pair around every kernel
 * on the handle's stream, so that wg_last_timing reports per-kernel durations;
 * at most 366 days per call. */
int wg_step_days_profiled(wg_handle *h, int32_t forcing_day0, int32_t first_doy, int32_t n_days);

/* Timing of the last wg_step_days / wg_step_days_profiled call (ms, HIP events on
 * the handle's stream): [0] cell kernels, [1] routing-chain kernels, [2] whole
 * call (only [2] for wg_step_days).  n_launches receives the launch counts. */
int wg_last_timing(wg_handle *h, double *ms, int64_t *n_launches);

/* Autotuner, no counterpart in the reference: times the launch configurations
 * (cell-kernel tile 128 or 256 cells, chain kernel mode 0, 1 or 2 with a
 * workgroup of 64, 128 or 256 threads or mode 3 with 256 or 512, snow chunk 5 or
 * 10 sub-cell pairs)
 * on n_days days of the loaded forcing, each from the same saved state and run
 * as wg_step_days runs them (graphs, partitions), keeps the fastest and restores
 * the state.  Needs device memory for one copy of the
 * handle's buffers.  config receives [cell threads, chain threads, snow chunk,
 * chain mode]; ms_per_day the best day time.  Results are identical under every
 * configuration. */
int wg_autotune(wg_handle *h, int32_t n_days, int32_t *config, double *ms_per_day);
/* Sets the launch configuration directly (cell threads 64..256 in steps of 64,
 * chain threads 64..256 (mode 3: 512 runs eight wavefronts, less four), snow
 * chunk 5 or 10 (the chunk-5 cell kernel is held to three wavefronts per SIMD,
 * which suits many members; the chunk-10 one is not), chain mode 0 level walk
 * loading one level ahead, 1 dataflow
 * walk without per-level barriers, 2 level walk loading two levels ahead, 3 wave
 * walk in batches of 64 cells, the default). */
int wg_set_launch_config(wg_handle *h, int32_t cell_threads, int32_t chain_threads, int32_t snow_chunk,
                         int32_t chain_mode);

/* Schedule facts: [0] routing groups, [1] deepest group (levels), [2] level
 * slots, [3] cells routed by the cell kernel (no upstream cell). */
int wg_schedule_info(wg_handle *h, int32_t *info);
/* No counterpart in the reference (its routing is one serial loop).  The routing
 * groups of the day loop fall into one or two partitions: with two (handles of at
 * most 8 members), the deepest groups (at least 0.8 times the deepest, at most 30 %
 * of the cells) run their days as a graph on a stream of their own beside the
 * others', which changes no result.
 * Returns the number of partitions; cells[k] and depth[k] (may be NULL) receive the
 * cells and deepest group (levels) of the first max_parts of them.  The environment
 * variable WG_PARTITIONS=1, read by wg_create, keeps one. */
int wg_partitions(wg_handle *h, int32_t max_parts, int32_t *cells, int32_t *depth);

/* Diagnostic with no reference counterpart: evaluates the library's FP64 math on
 * the device, fn 0 exp(x), 1 pow(x, y), 2 sqrt(x), 3 log(x), 4 sin(2 pi x),
 * 5 cos(2 pi x), for n arguments in host arrays.  Used to check that the device
 * and the host produce the same bits. */
int wg_libm_eval(int32_t fn, int32_t n, const double *x, const double *y, double *out);

/* ---- calibration: the per-basin search for the runoff coefficient gamma ---- */

/* State of one basin's search: the function-static variables of
 * calibGammaClass::findNewGamma (calibration.cpp:275-284) and the members it
 * updates (calibration.h:13-15, 46-48), one struct per basin so that basins can be
 * searched side by side. */
typedef struct wg_gamma_search {
  float gamma_low, gamma_high;   /* bracket, -99 = not yet found */
  float sum_diff_old;            /* sumOfDifferences of the previous call */
  int32_t call_counter;
  int32_t gamma_cond;            /* 1 found, 2 limit reached (10 % rule / CFA) */
  int32_t calib_status;          /* 1 gamma, 2 gamma + 10 % rule, 3 + CFA */
  int32_t cell_corr_factor_ind;  /* 99: correction factor computed */
  float cell_corr_factor;        /* basin-wide CFA */
  float gamma_upper_limit, gamma_lower_limit;  /* 5 and 0.1 (calibration.cpp:45-48) */
  float last_nash_sutcliffe, last_sum_diff, measured_avg_adapt;  /* CALIBRATION.OUT columns */
  float sim_runoff_avg;          /* simRunoffSum / numberOfYears of the last call */
} wg_gamma_search;

/* Fresh search state (calibGammaClass constructor + function statics). */
void wg_gamma_search_init(wg_gamma_search *s);
/* Start value of gamma from the parameter file (calibration.cpp:196-205). */
float wg_gamma_start(const wg_gamma_search *s, float gamma_from_file);
/* calibGammaClass::findNewGamma (calibration.cpp:266-636): compares the annual
 * simulated and observed discharge of the evaluation years (observed <= -1 = no
 * data) and returns the next gamma, or -99 when the search ends.  sim_inflow and
 * sim_water_use may be NULL (zero). */
float wg_find_new_gamma(wg_gamma_search *s, float gamma_old, int32_t n_years, const float *measured,
                        const float *simulated, const float *sim_inflow, const float *sim_water_use);

/* Station correction factor after a search (calibGammaClass::writeCorrFactors,
 * calibration.cpp:656-728): with the cell correction factor computed
 * (cell_corr_factor_ind 99) the measured sum over the years with data is scaled by
 * 1.1 (gamma at its upper limit) or 0.9 (lower limit) and CFS = measured / simulated;
 * within 1 % of 1 it is 1.  Sets calib_status 4 when CFS differs from 1. */
float wg_station_corr_factor(wg_gamma_search *s, float gamma, int32_t n_years, const float *measured,
                             const float *simulated);

/* calibGammaClass::createCorrectionGrid (calibration.cpp:730-801): the cell
 * correction factors of one calibration basin.  annual_pot[n_years][N] holds the
 * annual potential cell runoff of the evaluation years as the reference stores it
 * (G_POT_CELL_RUNOFF_<year>.UNF0, float).  Their mean (float, mean_pot when not NULL)
 * weighs the basin's discharge difference: for each cell with calib_basin[i] ==
 * station_number, corr[i] = 1 - sign(mean) * (simulated - measured) / sum |mean|,
 * limited to 0.5 .. 1.5; simulated is the mean simulated discharge of the years and
 * measured the adapted measured mean (findNewGamma's arguments, calibration.cpp:627,
 * wg_gamma_search.sim_runoff_avg and .measured_avg_adapt).  Other cells of corr keep
 * their values. */
int wg_cell_corr_factor_grid(int32_t n_cells, const int16_t *calib_basin, int32_t station_number, int32_t n_years,
                             const float *annual_pot, float simulated, float measured, float *mean_pot, double *corr);

/* Calibration basins from station cells (1-based) in the order of the station file
 * (cbasinClass::prepare with upstreamBasin, calib_basins.cpp:25-300, 302-378): each
 * station claims the cells upstream of it that hold its own cell's value (no basin
 * yet, or the earlier basin it lies in, which it then splits).  Station cells <= 1
 * (outside the land mask) and repeats are skipped (used[k] = 0).  calib_basin[N]
 * receives the basin number (1-based, in the order of the used stations), 0 outside
 * every basin, and -1 for the cell below a station that lies in no basin
 * (calib_basins.cpp:270-285).  overlap[k]: 0 none, 1 the station lies in an earlier
 * basin, 2 earlier basins lie upstream, 3 both (-1 when skipped).  used and overlap
 * may be NULL.  Returns the number of basins, or a negative WG_E_* code. */
int wg_calib_basins(int32_t n_cells, const int32_t *inflow9, const int32_t *outflow, int32_t n_stations,
                    const int32_t *station_cells, int32_t *used, int16_t *calib_basin, int32_t *overlap);
/* The stations upstream of each basin's station (upstreamStationClass::findStations,
 * upstream_stations.cpp:66-140): all_stations 0 stops at the first station on each
 * path (directUpstSt), 1 goes through (allUpstSt).  basin_cells are the basins'
 * station cells (1-based) in basin order.  Basin b's list is stations[offsets[b] ..
 * offsets[b+1]) (1-based basin numbers, in the reference's list order); WG_E_ARG when
 * more than max_entries entries are found. */
int wg_upstream_stations(int32_t n_cells, const int32_t *inflow9, int32_t n_basins, const int32_t *basin_cells,
                         int32_t all_stations, int32_t *offsets, int32_t *stations, int32_t max_entries);

/* ---- calibration surrogate: 16-128-128-128-1 GELU perceptron on BF16 MFMA ----
 * No counterpart in the reference: it proposes gamma from basin features so that
 * the plain model only has to verify (the search it replaces is
 * calibration.cpp:266-636).  Parameters are one flat FP32 vector: W1 [128][16],
 * b1 [128], W2 [128][128], b2, W3 [128][128], b3, w4 [128], b4; GEMMs run on
 * BF16 copies with FP32 accumulation. */
typedef struct wg_mlp wg_mlp;
int wg_mlp_create(uint64_t seed, wg_mlp **out);
int wg_mlp_destroy(wg_mlp *h);
const char *wg_mlp_last_error(const wg_mlp *h);
int64_t wg_mlp_n_params(void);
int wg_mlp_get_params(wg_mlp *h, float *out);
int wg_mlp_set_params(wg_mlp *h, const float *in);
/* y[r] for rows x[r][16] (host arrays). */
int wg_mlp_forward(wg_mlp *h, int32_t rows, const float *x, float *y);
/* wg_mlp_forward without any allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight proposal loops).  No counterpart in the
 * reference.  x [rows][16] and y [rows] are device pointers on the handle's
 * device; the work is queued on `stream` (a hipStream_t of that device; NULL: the
 * handle's own stream) and not waited for.  The activations live in one static
 * scratch array per GPU of WG_MLP_NOALLOC_MAX_ROWS rows (6.5 MiB, held until the
 * process ends): 1 <= rows <= WG_MLP_NOALLOC_MAX_ROWS, else WG_E_ARG.  Every call
 * on a GPU shares that array, so calls that overlap on one GPU (other streams,
 * other handles) race.  Results are wg_mlp_forward's bits. */
#define WG_MLP_NOALLOC_MAX_ROWS 8192
int wg_mlp_forward_noalloc(wg_mlp *h, int32_t rows, const float *x, float *y, void *stream);
/* Mean-squared-error loss and its gradient (flat, like the parameters) over rows. */
int wg_mlp_gradient(wg_mlp *h, int32_t rows, const float *x, const float *t, float *loss, float *grad);
/* Adam training over shuffled mini-batches (batch a multiple of 32); mean batch
 * loss per epoch into epoch_loss[epochs] (may be NULL). */
int wg_mlp_train(wg_mlp *h, int32_t n_samples, const float *x, const float *t, int32_t epochs, int32_t batch,
                 float lr, uint64_t seed, float *epoch_loss);

#ifdef __cplusplus
}
#endif
#endif
