// Closed-shell V2 blocks from Cholesky vectors on the device: the GEMMs over
// the Cholesky index of V2TensorSetup::setupV2Tensors (cholesky/v2tensors.cpp:52-98),
// in the spatial closed-shell form this library consumes, (pq|rs) =
// sum_Q L[p][q][Q] L[r][s][Q].  The spin-orbital <pq||rs> of the reference
// blocks ijab/ijka/iabc are formed from these spatial integrals wherever they
// are read (xt_kernels.hip, xt_gather.hip).  Plain DGEMMs go to rocBLAS.
//
// L is [nmo][nmo][K] with nmo = O + V (occupied orbitals first) and the
// Cholesky index fastest, the order of cholVpr(p, q, cind).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "xt_device.h"

namespace xt {

namespace {
struct Freer {
  void* p[4] = {nullptr, nullptr, nullptr, nullptr};
  ~Freer() {
    for(void* x: p)
      if(x) track_free(x);
  }
};
int blas_err(rocblas_status s) { return s == rocblas_status_success ? 0 : (int) hipErrorUnknown; }
}  // namespace

int cholesky_integrals(int O, int V, long K, const double* L, double* vvoo, double* ovoo,
                       double* xv, void* stream) {
  hipStream_t  st = (hipStream_t) stream;
  const long   o = O, v = V, nmo = o + v, vp = (V + 7) / 8 * 8;
  const size_t d = sizeof(double);
  Freer        f;
  double *     lvo = nullptr, *lov = nullptr, *loo = nullptr, *tmp = nullptr;
  hipError_t   e;
  if((e = (hipError_t) track_malloc((void**) &lvo, d * v * o * K))) return (int) e;
  f.p[0] = lvo;
  if((e = (hipError_t) track_malloc((void**) &lov, d * o * v * K))) return (int) e;
  f.p[1] = lov;
  if((e = (hipError_t) track_malloc((void**) &loo, d * o * o * K))) return (int) e;
  f.p[2] = loo;
  // compact row sets: L_vo[(a,i)] = L[O+a][i], L_ov[(i,a)] = L[i][O+a], L_oo[(i,j)] = L[i][j]
  if((e = hipMemcpy2DAsync(lvo, d * o * K, L + o * nmo * K, d * nmo * K, d * o * K, v,
                           hipMemcpyDeviceToDevice, st)))
    return (int) e;
  if((e = hipMemcpy2DAsync(lov, d * v * K, L + o * K, d * nmo * K, d * v * K, o,
                           hipMemcpyDeviceToDevice, st)))
    return (int) e;
  if((e = hipMemcpy2DAsync(loo, d * o * K, L, d * nmo * K, d * o * K, o, hipMemcpyDeviceToDevice,
                           st)))
    return (int) e;
  rocblas_handle h = nullptr;
  if(rocblas_create_handle(&h) != rocblas_status_success) return (int) hipErrorUnknown;
  struct HandleGuard {
    rocblas_handle h;
    ~HandleGuard() { (void) rocblas_destroy_handle(h); }
  } hg{h};
  int rc = blas_err(rocblas_set_stream(h, st));
  if(rc) return rc;
  const double one = 1.0, zero = 0.0;
  const int    ik = (int) K;
  // Column-major views: a row-major [rows][K] matrix is a K x rows column-major
  // matrix with leading dimension K, so every product below is op = (T, N).
  // ovoo[i][a][j][k] = (ia|jk) = sum_Q L_ov[(i,a)][Q] L_oo[(j,k)][Q]
  rc = blas_err(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (int) (o * o),
                              (int) (o * v), ik, &one, loo, ik, lov, ik, &zero, ovoo, (int) (o * o)));
  if(rc) return rc;
  // (ai|bj) over rows (a,i), columns (b,j), then into vvoo[a][b][i][j]
  if((e = (hipError_t) track_malloc((void**) &tmp, d * v * o * v * o))) return (int) e;
  f.p[3] = tmp;
  rc = blas_err(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (int) (v * o),
                              (int) (v * o), ik, &one, lvo, ik, lvo, ik, &zero, tmp, (int) (v * o)));
  if(rc) return rc;
  if((rc = launch_vvoo_order(tmp, vvoo, O, V, st))) return rc;
  // xv[c][b][i][a] = vvvo[a][c][b][i] = (ac|bi) = sum_Q L[O+a][O+c][Q] L_vo[(b,i)][Q]:
  // for each c a (V x K) by (K x V*O) product, the a operand rows strided by
  // nmo*K and the batch by K; the a-padding of xv stays zero
  if((e = hipMemsetAsync(xv, 0, d * v * v * o * vp, st))) return (int) e;
  rc = blas_err(rocblas_dgemm_strided_batched(
    h, rocblas_operation_transpose, rocblas_operation_none, (int) v, (int) (v * o), ik, &one,
    L + (o * nmo + o) * K, (int) (nmo * K), K, lvo, ik, 0, &zero, xv, (int) vp, v * o * vp, (int) v));
  if(rc) return rc;
  return (int) hipStreamSynchronize(st);
}

}  // namespace xt
