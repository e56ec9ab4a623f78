// Six-tile halo update as one gather launch.
//
// Reference: the halo exchanges of fv3gfs-util's CubedSphereCommunicator
// (scalar and vector updaters built in fv3core/stencils/dyn_core.py:216-286).
// With every tile in one batched storage, an update is a gather through an
// index map built on the host (fv3core_amd/halo.py): for each halo point,
// the source tile, source indices, source component and sign.  One thread
// handles one (map entry, level) and moves that value for every field group
// of the call.  A batch of several cubes (ensemble members, six consecutive
// tiles each) applies the same six-tile map to every member: grid.y is the
// member and offsets both tile indices by six per member.
#include "common.hpp"

namespace fv3 {

struct FieldPtrs {
  void* p[16];
};

template <typename T>
__global__ void __launch_bounds__(256)
    halo_kernel(Dom d, int nent, const int* __restrict__ mp, int ngroups, int ncomp,
                FieldPtrs fp) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)nent * d.nk) return;
  const int e = (int)(n % nent);
  const int k = (int)(n / nent);
  const int* r = mp + 8 * (int64_t)e;
  const int m6 = 6 * (int)blockIdx.y;  // first tile of this member
  const int t = m6 + r[0], i = r[1], j = r[2], c = r[3];
  const int tb = m6 + r[4], ib = r[5], jb = r[6], sc = r[7];
  const int src_comp = sc >> 1;
  const bool neg = (sc & 1) != 0;
  const int64_t od = (int64_t)t * d.st + (int64_t)k * d.sk + (int64_t)j * d.sj + i;
  const int64_t os = (int64_t)tb * d.st + (int64_t)k * d.sk + (int64_t)jb * d.sj + ib;
  for (int g = 0; g < ngroups; ++g) {
    const T* src = reinterpret_cast<const T*>(fp.p[g * ncomp + src_comp]);
    T* dst = reinterpret_cast<T*>(fp.p[g * ncomp + c]);
    const T v = src[os];
    dst[od] = neg ? -v : v;
  }
}

template <typename T>
static int halo_launch(const fv3_domain* dom, int nent, const int* mp, int nfields, int ncomp,
                       const FieldPtrs* fp, void* stream) {
  const char* who = "fv3_halo_update";
  int rc = check_domain(dom, who);
  if (rc) return rc;
  FV3_REQUIRE(mp && fp, FV3_E_ARG, "%s: NULL pointer", who);
  FV3_REQUIRE(ncomp == 1 || ncomp == 2, FV3_E_ARG, "%s: ncomp must be 1 or 2", who);
  FV3_REQUIRE(nfields >= 1 && nfields <= 16 && nfields % ncomp == 0, FV3_E_ARG,
              "%s: bad field count %d", who, nfields);
  FV3_REQUIRE(dom->ntiles >= 6 && dom->ntiles % 6 == 0 && dom->ntiles / 6 <= 65535, FV3_E_ARG,
              "%s: the batch must hold whole six-tile cubes (got %d tiles)", who, dom->ntiles);
  for (int a = 0; a < nfields; ++a)
    FV3_REQUIRE(fp->p[a] != nullptr, FV3_E_ARG, "%s: field %d is NULL", who, a);
  if (nent == 0 || dom->nk == 0) return FV3_OK;
  Dom d = to_dom(dom);
  const int64_t total = (int64_t)nent * d.nk;
  hipLaunchKernelGGL((halo_kernel<T>), dim3((unsigned)((total + 255) / 256), d.ntiles / 6),
                     dim3(256), 0, as_stream(stream), d, nent, mp, nfields / ncomp, ncomp, *fp);
  FV3_HIP_CHECK_LAUNCH(who);
  return FV3_OK;
}

}  // namespace fv3

extern "C" {

int32_t fv3_halo_update_f64(const fv3_domain* d, int32_t nentries, const int32_t* map,
                            int32_t nfields, int32_t ncomp, const void* fields,
                            void* stream) {
  return fv3::halo_launch<double>(d, nentries, map, nfields, ncomp,
                                  reinterpret_cast<const fv3::FieldPtrs*>(fields), stream);
}

int32_t fv3_halo_update_f32(const fv3_domain* d, int32_t nentries, const int32_t* map,
                            int32_t nfields, int32_t ncomp, const void* fields,
                            void* stream) {
  return fv3::halo_launch<float>(d, nentries, map, nfields, ncomp,
                                 reinterpret_cast<const fv3::FieldPtrs*>(fields), stream);
}
}
