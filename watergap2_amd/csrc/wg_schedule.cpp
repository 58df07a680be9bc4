// wg_schedule.cpp -- routing schedule on the host: inflow/outflow cells from the
// LDD and the routing order with its Kahn sweeps.
//
// wg_inflow_outflow follows rout_prepare.cpp:338-439 (inflow_cells) and 619-670
// (outflow_cell); wg_rout_order follows rout_prepare.cpp:834-886 (rout_order).
// The sweep loop is the reference's algorithm (the order numbers it assigns are
// the interface), written as a frontier queue: a sweep lists the cells whose
// remaining inflow count is zero in ascending cell index, numbers them, then
// releases their outflow cells.  This touches each cell O(1) times instead of
// scanning all cells per sweep, and gives identical numbers.
#include <algorithm>
#include <vector>

#include "../../include/watergap2_amd.h"

namespace {

inline int32_t lattice_cell(const int32_t *cell_at, int col, int row) {
  if (row < 1 || row > 360) return 0;  // GCRC_2 border rows (rout_prepare.cpp:132-146)
  if (col < 1) col += 720;            // longitude wrap
  if (col > 720) col -= 720;
  return cell_at[(row - 1) * 720 + (col - 1)];
}

// D8 keypad directions: LDD code -> (dcol, drow), row 1 is north.
struct Dir {
  int dc, dr;
};
constexpr Dir kDir[10] = {{0, 0},  {-1, 1}, {0, 1},   {1, 1},  {-1, 0},
                          {0, 0},  {1, 0},  {-1, -1}, {0, -1}, {1, -1}};

}  // namespace

extern "C" int wg_inflow_outflow(int32_t n, const int8_t *ldd, const int16_t *row, const int16_t *col,
                                 const int32_t *cell_at, int32_t *inflow9, int32_t *outflow) {
  if (n < 0 || !ldd || !row || !col || !cell_at || !inflow9 || !outflow) return WG_E_ARG;
  // Channel of the inflow slot for a neighbour that flows in with LDD code c
  // (rout_prepare.cpp:363-418): the slot is the keypad position of the neighbour
  // seen from the receiving cell, minus one.
  static const int kSlot[10] = {-1, 8, 7, 6, 5, -1, 3, 2, 1, 0};
  for (int32_t i = 0; i < n; i++) {
    for (int k = 0; k < 9; k++) inflow9[(int64_t)i * 9 + k] = 0;
  }
  for (int32_t i = 0; i < n; i++) {
    const int c = ldd[i];
    outflow[i] = 0;
    if (c >= 1 && c <= 9 && c != 5) {
      const int32_t g = lattice_cell(cell_at, col[i] + kDir[c].dc, row[i] + kDir[c].dr);
      outflow[i] = g;
      // The receiving cell lists this cell in the slot of the neighbour position.
      if (g > 0) inflow9[(int64_t)(g - 1) * 9 + kSlot[c]] = i + 1;
    }
  }
  return WG_OK;
}

extern "C" int wg_rout_order(int32_t n, const int32_t *inflow9, const int32_t *outflow,
                             int32_t *rout_order, int32_t *level) {
  if (n < 0 || !inflow9 || !outflow || !rout_order) return WG_E_ARG;
  std::vector<int32_t> pending(n, 0);
  for (int32_t i = 0; i < n; i++) {
    rout_order[i] = -99;
    for (int k = 0; k < 9; k++)
      if (inflow9[(int64_t)i * 9 + k] > 0) pending[i]++;
  }
  for (int32_t i = 0; i < n; i++)
    if (outflow[i] < 0 || outflow[i] > n) return WG_E_ARG;
  std::vector<int32_t> frontier, next;
  for (int32_t i = 0; i < n; i++)
    if (pending[i] == 0) frontier.push_back(i);
  int32_t number = 1, sweep = 0, done = 0;
  while (!frontier.empty()) {
    // cells of one sweep are numbered in ascending cell index (rout_prepare.cpp:859-863)
    std::sort(frontier.begin(), frontier.end());
    for (int32_t c : frontier) {
      rout_order[c] = number++;
      if (level) level[c] = sweep;
      pending[c] = -1;
      done++;
    }
    // Release the outflow cells; a cell joins the next sweep only if its count is
    // exactly zero after all releases of this sweep.  A count driven below zero
    // (an outflow cell that does not list this cell as inflow) is never numbered,
    // as in the reference, where such a counter never returns to zero.
    next.clear();
    for (int32_t c : frontier) {
      const int32_t d = outflow[c];
      if (d > 0 && --pending[d - 1] == 0) next.push_back(d - 1);
    }
    frontier.clear();
    for (int32_t c : next)
      if (pending[c] == 0) frontier.push_back(c);
    sweep++;
  }
  return done == n ? WG_OK : WG_E_CYCLE;
}
