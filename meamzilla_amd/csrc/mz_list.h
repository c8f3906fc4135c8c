// Device-side pair-list build (see mz_list.hip).
#pragma once

#include <string>

struct mz_ctx;

namespace mz {

// Builds c->atom_poff and c->d_atom_poff / c->d_pairs (and, with host_copy,
// c->pairs) for the cells, types, spline grids and radial cutoff already set
// in c; also uploads positions, boxes, types and the spline end points.
// Returns MZ_OK or an MZ_ERR_* code with the message in err.
int build_pairs_device(mz_ctx* c, const double* pos, const double* box, double max_rcut, bool host_copy,
                       std::string& err);

// Builds c->in_off and c->d_in_off / c->d_in_list (the pairs whose neighbour
// is each atom, ascending) from c->d_pairs.
int build_incoming_device(mz_ctx* c, std::string& err);

}  // namespace mz
