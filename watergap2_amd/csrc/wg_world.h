/*
 * wg_world.h -- seeded synthetic WaterGAP grids (host C++).
 *
 * The reference ships no input data, so every workload is a synthetic 0.5-degree
 * grid built here: a fractal DEM, priority-flood pit filling that also assigns the
 * D8 flow direction of each cell, the routing schedule, static fields with the
 * reference's element types, and daily forcing.  All randomness comes from
 * Philox4x32-10 keyed by (seed, stream) with a per-cell/per-day counter, so any
 * sub-range of days or cells can be regenerated independently.
 */
#ifndef WG_WORLD_H
#define WG_WORLD_H

#include <stdint.h>
#include "../../include/watergap2_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wg_world_spec {
  int32_t kind;          /* 0: regional lattice, all land, single outlet; 1: global land mask */
  int32_t nrows, ncols;  /* kind 0 lattice size */
  int32_t row0, col0;    /* kind 0: 1-based global row/column of the lattice's north-west cell */
  int32_t n_land;        /* kind 1: number of land cells */
  int32_t n_outlets;     /* kind 1: number of coastal outlet seeds */
  int32_t n_sinks;       /* kind 1: number of inland sinks (LDD -1) */
  uint64_t seed;
  double water_frac_max; /* local lake / local wetland / global wetland % drawn from U(0, 100*max) */
  int32_t glolake_every; /* one global lake per this many cells (0: none) */
  int32_t res_every;     /* one reservoir per this many cells (0: none) */
  int32_t edge_cases;    /* 1: add cells that exercise rarely taken branches */
  double relief_scale;   /* kind 1: length scale (cells) of the relief's largest octave; 0 = 40 */
} wg_world_spec;

typedef struct wg_world {
  int32_t n;
  int16_t *row, *col;
  int32_t *cell_at;      /* [360*720] 1-based cell at each position, 0 = no land */
  double *cell_area, *contfreq;
  int16_t *aindex;
  int8_t *ldd, *land_cover, *texture, *res_type, *start_month;
  float *built_up, *smax, *gw_factor;
  int16_t *rgmax;
  int16_t *elevation;    /* [N][101] */
  double *params;        /* [WG_NPARAM][N] */
  double *glo_lake, *loc_lake, *glo_wetland, *loc_wetland, *lake_area, *reservoir_area, *glo_res;
  double *stor_cap, *mean_outflow;
  float *river_length, *river_slope, *roughness, *bankfull_flow, *altitude, *meander;
  int32_t *inflow9, *downstream, *rout_order, *level;
  int32_t *basin;        /* 0-based basin id of each cell */
  int32_t n_basins;
  int32_t *basin_outlet; /* [n_basins] 0-based outlet cell of each basin */
  int32_t max_level;
  double *upstream_area; /* km2 including the cell */
  wg_lct_tables lct;
  int32_t kind;
  uint64_t seed;
} wg_world;

int wg_world_create(const wg_world_spec *spec, wg_world **out);
void wg_world_destroy(wg_world *w);
/* Fills the pointer struct of wg_create from a world. */
void wg_world_statics(const wg_world *w, wg_static_fields *out);
/* Daily forcing [n_days][4][N] (T degC, P mm/d, SW W/m2, LW W/m2) for global day
 * indices day0..day0+n_days-1 (julian day = index % 365 + 1). */
int wg_world_forcing(const wg_world *w, uint64_t seed, int64_t day0, int32_t n_days, float *out);

/* Sub-world of the cells of a basin set (cells keep their relative order),
 * used to cut calibration basins from a global grid.  cells are 0-based. */
int wg_world_subset(const wg_world *w, int32_t n_cells, const int32_t *cells, wg_world **out);

#ifdef __cplusplus
}
#endif
#endif
