// Plain-old-data tables shared by the host planner and the HIP kernels.
#pragma once

#include <stdint.h>

#include "fbme_hip.h"

// Per-scene descriptor of a (possibly uneven) batch.  Scenes are stored back
// to back in every buffer; the offsets below are element offsets.
//   grid buffers   [N1*N2]   gamma, mean, variance, backprojection, score, x
//   sample buffers [M1*M2]   data, message, prediction, observation mask
//                            (one byte per sample; unobserved samples carry
//                            zero messages and no residual)
//   mid buffers    [N1*M2]   column-pass <-> row-pass intermediate
struct FbmeScene {
  int n1, n2, m1, m2;
  unsigned rad1, rad2;      // radix codes for FFT lengths n1 and n2
  int iters;                // UAMP iteration budget T_u
  int k;                    // active-set capacity min(polish_size, N)
  long long grid_off;
  long long samp_off;
  long long mid_off;
  int col_items, row_items; // work items per scene in column / row passes
  int col_part, row_part;   // first partial slot of this scene
  int cand_off;             // first candidate slot (row_items * k per scene)
  int dir1_off, dir2_off;   // Dirichlet tables (double2) for the Gram kernel
  int xtile_items;          // exchange-scoring / output tiles
  int xcand_off;            // first exchange candidate slot
  int orig;                 // caller's scene index (per-scene outputs)
  int tw1_off, tw2_off;     // twiddle tables (float2) of lengths n1 and n2
  int m_obs;                // observed samples (M1 * M2 without a mask)
  int mbits_off;            // register-pass mask bits (uint16 [M2*16]) or -1
  long long mh_off;         // mask spectrum (double2 [N1*N2]) or -1 without a mask
};

// Kernel variant carried in the `ept` argument of the pass launchers: 4, 8 or
// 16 select the LDS Stockham kernels with that many elements per thread;
// FBME_EPT_R16 selects the register-resident 16 x 16 FFT kernels (transform
// length 256 along the pass axis), FBME_R16_COLS columns per column workgroup.
#define FBME_EPT_R16 256
#define FBME_R16_COLS 16
#define FBME_R16_ROWS 4

// One unit of work of a pass: `count` consecutive lines (columns for the
// column passes, rows for the row passes) of scene `scene`, starting at
// `first`.  `slot` is the scene-local partial-sum slot.
struct FbmeWork {
  int scene, first, count, slot;
};

// Per-scene scalar state (doubles).  Written by the scalar kernels and read
// by the passes; float32 state is stored rounded to float.
enum FbmeScalar {
  SC_POWER = 0,     // mean |y|^2 (clamped at eps)
  SC_FLOOR = 1,     // variance floor max(1e-12 P, tiny)
  SC_NOISE = 2,     // sigma^2
  SC_DV = 3,        // data variance sum(v)
  SC_DOF = 4,       // effective dof (clamped)
  SC_BPE = 5,       // sum |A^H y|^2
  SC_YE = 6,        // ||y||^2
  SC_GSUM = 7,      // sum gamma after the sweep (diffuse tail loading only)
  SC_CHANGE = 8,    // ||mu_t - mu_t-1|| / max(||mu_t||, eps) of the last tracked step
  SC_COUNT = 9
};

// Per-scene partial sums (doubles), `FBME_PART_STRIDE` slots per quantity.
enum FbmePart {
  PT_YE = 0,   // sum |y|^2                      (column pass, init)
  PT_DOF = 1,  // sum relevance                  (row pass)
  PT_VS = 2,   // sum posterior variance         (row pass)
  PT_RES = 3,  // sum |y - A mu|^2               (column pass)
  PT_DMU = 4,  // sum |mu_t - mu_t-1|^2          (row pass, tracked steps)
  PT_MU = 5,   // sum |mu_t|^2                   (row pass, tracked steps)
  PT_COUNT = 6
};



#define FBME_KMAX 32
#define FBME_TILE_MAX 4096     // complex elements per FFT tile (EPT * 256)
#define FBME_XTILE 1024        // grid points per exchange-scoring tile
