/* The reference's time loop driven through the C ABI from plain C99 (the
 * loop of ActFlow_V3-2.cu:718-862 with af_run_step in chunks), as a caller
 * would write it:
 *   step_loop <N> <precision> <steps> <chunk> <snapshot path>
 * prints one line per step, "<step> <energy %.17g>", then "M <Mx> <My>", and
 * writes the final state as a snapshot file. */
#include <stdio.h>
#include <stdlib.h>

#include "actflow.h"

int main(int argc, char** argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: step_loop N precision steps chunk snapshot\n");
    return 2;
  }
  const int n = atoi(argv[1]), prec = atoi(argv[2]), steps = atoi(argv[3]), chunk = atoi(argv[4]);
  af_params p;
  af_params_jw(&p, n, n, 6.283185307179586, 6.283185307179586, 2e-4, 3.5, -0.8, 0.01, -0.045, 2.5e-5);
  p.precision = prec;
  const size_t nd = 2 * (size_t)n * (size_t)(n / 2 + 1);
  double* w = malloc(nd * sizeof(double));
  double* e = malloc((size_t)chunk * sizeof(double));
  double M[2] = {0.05 * n * n, -0.02 * n * n};
  af_run* run = NULL;
  int rc = 1;
  if (!w || !e) goto done;
  if (af_ic_bandlimited(n, n, 11, 1.0, 8.0, w) != AF_OK) goto fail;
  if (af_run_create(&p, &run) != AF_OK) goto fail;
  if (af_run_set_state(run, w, M) != AF_OK) goto fail;
  for (int m = 0; m < steps; m += chunk) {
    const int k = steps - m < chunk ? steps - m : chunk;
    long long nan_step = -1;
    const af_status st = af_run_step(run, k, e, &nan_step);
    if (st == AF_ENAN) {
      fprintf(stderr, "array contains nan values after %lld timesteps\n", nan_step);
      goto done;
    }
    if (st != AF_OK) goto fail;
    for (int i = 0; i < k; ++i) printf("%d %.17g\n", m + i + 1, e[i]);
  }
  if (af_run_get_state(run, w, M) != AF_OK) goto fail;
  if (af_snapshot_write(argv[5], n, n, w, M, 17) != AF_OK) goto fail;
  printf("M %.17g %.17g\n", M[0], M[1]);
  rc = 0;
  goto done;
fail:
  fprintf(stderr, "error: %s\n", af_last_error());
done:
  af_run_destroy(run);
  free(w);
  free(e);
  return rc;
}
