// Launchers shared between translation units of the library.
#pragma once

#include "stencil.hpp"

namespace fv3 {

// AGrid2BGridFourthOrder.__call__ (a2b_ord4.py:617-695) on levels
// [k0, k0 + nk): qout on the corners [isc, iec+1]^2; qx, qy and edges are
// scratch slabs laid out like the fields.  With replace set, qin_w (the
// writable qin) receives qout as well.
template <typename T>
int launch_a2b_ord4(hipStream_t s, const fv3_domain* dom, const Met<T>& m, int k0, int nk,
                    const T* qin, T* qout, T* qx, T* qy, T* edges, int replace, T* qin_w,
                    const char* who);

int check_a2b_metrics(const fv3_metrics* g, const char* who);

}  // namespace fv3
