/* A plain C client of include/quantit_amd.h: what a foreign-function binding
 * (cgo, JNI, ctypes) does through the C ABI.  Two U(1) block-diagonal
 * matrices are created from host arrays, contracted on the device and read
 * back; the product is checked block by block against a host product, and an
 * invalid block index must come back as QTT_INVALID_ARGUMENT with a message.
 * Exit status 0 on success. */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "quantit_amd.h"

#define CHECK(call)                                                              \
  do {                                                                           \
    int st_ = (call);                                                            \
    if (st_ != QTT_OK) {                                                         \
      fprintf(stderr, "%s failed: %d %s\n", #call, st_, qtt_last_error());       \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

int main(void) {
  qtt_context* ctx = NULL;
  CHECK(qtt_context_create(0, (size_t)64 << 20, &ctx));
  /* legs: rows {2 x charge 0, 3 x charge 1}, columns {2 x charge 0, 3 x charge -1} */
  const int64_t nsec[2] = {2, 2};
  const int64_t sizes[4] = {2, 3, 2, 3};
  /* A's column charges (0, -1) are the inverses of B's row charges (0, 1), so
     A and B can share one charge list */
  const int16_t cv[4] = {0, 1, 0, -1};
  const int16_t sel = 0;
  const uint16_t mod = 0;
  const int32_t idx[4] = {0, 0, 1, 1};
  double a[13], b[13];
  for (int i = 0; i < 13; ++i) {
    a[i] = 0.5 + i;
    b[i] = 1.0 / (1.0 + i);
  }
  qtt_btensor *A = NULL, *B = NULL, *Cm = NULL;
  CHECK(qtt_btensor_create(ctx, 2, nsec, sizes, 1, &mod, cv, &sel, 2, idx, a, &A));
  CHECK(qtt_btensor_create(ctx, 2, nsec, sizes, 1, &mod, cv, &sel, 2, idx, b, &B));
  const int64_t da = 1, db = 0;
  CHECK(qtt_tensordot(A, B, 1, &da, &db, &Cm));
  int rank = 0, ng = 0;
  int64_t ns = 0, nb = 0, ne = 0;
  CHECK(qtt_btensor_info(Cm, &rank, &ng, &ns, &nb, &ne));
  if (rank != 2 || nb != 2 || ne != 13) {
    fprintf(stderr, "unexpected result shape: rank %d, %lld blocks, %lld elements\n", rank, (long long)nb,
            (long long)ne);
    return 1;
  }
  int32_t cidx[4];
  double c[13];
  CHECK(qtt_btensor_blocks(Cm, cidx, c));
  /* host product block by block (row-major n x n blocks at offsets 0 and 4) */
  const int off[2] = {0, 4}, n[2] = {2, 3};
  double err = 0.0;
  for (int blk = 0; blk < 2; ++blk)
    for (int i = 0; i < n[blk]; ++i)
      for (int j = 0; j < n[blk]; ++j) {
        double s = 0.0;
        for (int k = 0; k < n[blk]; ++k) s += a[off[blk] + i * n[blk] + k] * b[off[blk] + k * n[blk] + j];
        const double d = fabs(s - c[off[blk] + i * n[blk] + j]);
        if (d > err) err = d;
      }
  if (err > 1e-12) {
    fprintf(stderr, "product differs by %g\n", err);
    return 1;
  }
  /* an index outside the sections is rejected before any device work */
  const int32_t bad[2] = {0, 5};
  qtt_btensor* X = NULL;
  const int st = qtt_btensor_create(ctx, 2, nsec, sizes, 1, &mod, cv, &sel, 1, bad, a, &X);
  if (st != QTT_INVALID_ARGUMENT || strlen(qtt_last_error()) == 0) {
    fprintf(stderr, "bad index: status %d\n", st);
    return 1;
  }
  CHECK(qtt_btensor_destroy(Cm));
  CHECK(qtt_btensor_destroy(B));
  CHECK(qtt_btensor_destroy(A));
  CHECK(qtt_context_destroy(ctx));
  printf("abi client ok: 2 blocks, max error %g\n", err);
  return 0;
}
