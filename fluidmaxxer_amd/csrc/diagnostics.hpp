// diagnostics.hpp -- the FP64 state sums of diagnostics.hip, for the fused run's periodic log.
#pragma once

#include <cstdint>

#include "fmx_internal.hpp"

namespace fmx {

// columns of one sum row: sum phi, sum q, sum of the stored populations f (NS density), non-finite count
constexpr int kMassCols = 4;
// the partial-sum grid: one block per 4096 cells, at most 2048 blocks (so at most 64 KiB of partials)
constexpr i64 kMassCellsPerBlock = 4096;
constexpr int kMassMaxBlocks = 2048;

int mass_blocks(i64 n);
// sums of the nullable phi / q (compute precision of prec) and f (storage precision) over n cells
// into out[0..3], partials in `partial` (mass_blocks(n) * kMassCols doubles)
int launch_mass_any(int prec, const void *phi, const void *q, const void *f, i64 n, double fref, double *partial,
                    double *out, hipStream_t s);

// FP64 sum of the stored populations over the cells without the solid link bit (out[0]) and their
// number (out[1]); partials as above, out holds kMassCols doubles
int launch_fluid_rho_any(int prec, const void *f, const std::uint32_t *links, i64 n, double fref, double *partial,
                         double *out, hipStream_t s);

}  // namespace fmx
