/*
 * A C caller of libsdegpu: the drop-in boundary exercised without Python.
 *
 * 1. Deterministic lognormal paths (sigma = 0): log-Euler telescopes, so the
 *    stored values must equal x0 * exp(mu * t) to rounding (the exactness of
 *    sdepy's lognorm_process, tests/test_processes.py:681-708).
 * 2. The same with sigma > 0: the FP64 cross-path mean of x_T from
 *    sdegpu_path_moments must lie within 4 standard errors of
 *    x0 * exp(mu * T) (analytical lognorm_mean, analytical.py:133-146).
 * 3. An invalid problem returns a status code and a message.
 * Exit status 0 on success.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "sdegpu.h"

#define CHECK_HIP(x)                                                          \
  do {                                                                        \
    if ((x) != hipSuccess) {                                                  \
      fprintf(stderr, "HIP error at %s:%d\n", __FILE__, __LINE__);            \
      return 2;                                                               \
    }                                                                         \
  } while (0)

static int run(double sigma, long long paths, int nsteps, double T, double* xT_host,
               double* mean, double* var) {
  const double x0 = 1.5, mu = 0.07;
  const int ncoef = 4; /* ds, sqrt ds, mu - sigma^2/2, sigma */
  double* tab = (double*)malloc(sizeof(double) * nsteps * ncoef);
  int* rows = (int*)malloc(sizeof(int) * nsteps);
  for (int s = 0; s < nsteps; ++s) {
    const double ds = T / nsteps;
    tab[s * ncoef + 0] = ds;
    tab[s * ncoef + 1] = sqrt(ds);
    tab[s * ncoef + 2] = mu - sigma * sigma / 2;
    tab[s * ncoef + 3] = sigma;
    rows[s] = s + 1; /* store every step: rows 1..nsteps */
  }
  void *d_tab, *d_rows, *d_out, *d_mean, *d_m2, *d_ws;
  const size_t out_bytes = sizeof(double) * (size_t)(nsteps + 1) * (size_t)paths;
  CHECK_HIP(hipMalloc(&d_tab, sizeof(double) * nsteps * ncoef));
  CHECK_HIP(hipMalloc(&d_rows, sizeof(int) * nsteps));
  CHECK_HIP(hipMalloc(&d_out, out_bytes));
  CHECK_HIP(hipMemcpy(d_tab, tab, sizeof(double) * nsteps * ncoef, hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d_rows, rows, sizeof(int) * nsteps, hipMemcpyHostToDevice));

  sdegpu_euler_problem p;
  memset(&p, 0, sizeof p);
  p.model = SDEGPU_MODEL_LOGNORM;
  p.dtype = SDEGPU_F64;
  p.nvars = 1;
  p.nout = 1;
  p.out_mode = SDEGPU_OUT_IDENTITY;
  p.exp_mask = 1u;
  p.paths = paths;
  p.nsteps = nsteps;
  p.ncoef = ncoef;
  p.coef = d_tab;
  p.store_row = (const int32_t*)d_rows;
  p.row0 = 0;
  p.x0[0] = log(x0);
  p.out = d_out;
  p.key_w[0] = 0x1234u;
  p.key_w[1] = 0x5678u;
  int rc = sdegpu_euler(&p, NULL);
  if (rc != SDEGPU_OK) {
    char msg[512];
    sdegpu_last_error(msg, sizeof msg);
    fprintf(stderr, "sdegpu_euler failed: %d %s\n", rc, msg);
    return 3;
  }
  /* moments of the last row */
  const size_t ws_bytes = sdegpu_moments_workspace(1, paths);
  CHECK_HIP(hipMalloc(&d_mean, sizeof(double)));
  CHECK_HIP(hipMalloc(&d_m2, sizeof(double)));
  CHECK_HIP(hipMalloc(&d_ws, ws_bytes));
  const double* last = (const double*)d_out + (size_t)nsteps * (size_t)paths;
  rc = sdegpu_path_moments(last, SDEGPU_F64, 1, paths, (double*)d_mean, (double*)d_m2, d_ws,
                           ws_bytes, NULL);
  if (rc != SDEGPU_OK) return 4;
  CHECK_HIP(hipDeviceSynchronize());
  double m2;
  CHECK_HIP(hipMemcpy(mean, d_mean, sizeof(double), hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemcpy(&m2, d_m2, sizeof(double), hipMemcpyDeviceToHost));
  *var = m2 / (double)paths;
  CHECK_HIP(hipMemcpy(xT_host, last, sizeof(double) * 4, hipMemcpyDeviceToHost));
  hipFree(d_tab);
  hipFree(d_rows);
  hipFree(d_out);
  hipFree(d_mean);
  hipFree(d_m2);
  hipFree(d_ws);
  free(tab);
  free(rows);
  return 0;
}

int main(void) {
  const double x0 = 1.5, mu = 0.07, T = 2.0;
  double xT[4], mean, var;
  int count = 0;
  if (sdegpu_device_count(&count) != SDEGPU_OK || count < 1) {
    fprintf(stderr, "no GPU\n");
    return 5;
  }
  printf("library version %d, %d device(s)\n", sdegpu_version(), count);

  int rc = run(0.0, 1000, 50, T, xT, &mean, &var);
  if (rc) return rc;
  const double exact = x0 * exp(mu * T);
  for (int i = 0; i < 4; ++i) {
    if (fabs(xT[i] - exact) > 1e-13 * exact) {
      fprintf(stderr, "sigma=0: x_T = %.17g, expected %.17g\n", xT[i], exact);
      return 6;
    }
  }
  printf("sigma=0: x_T = %.15f (exact %.15f)\n", xT[0], exact);

  const long long paths = 200000;
  rc = run(0.3, paths, 100, T, xT, &mean, &var);
  if (rc) return rc;
  const double se = sqrt(var / (double)(paths - 1));
  printf("sigma=0.3: mean x_T = %.6f +- %.6f (lognorm_mean %.6f)\n", mean, se, exact);
  if (fabs(mean - exact) > 4 * se) return 7;

  sdegpu_euler_problem bad;
  memset(&bad, 0, sizeof bad);
  bad.model = SDEGPU_MODEL_LOGNORM;
  bad.dtype = SDEGPU_F64;
  bad.nvars = SDEGPU_MAX_COUPLED + 1;
  rc = sdegpu_euler(&bad, NULL);
  char msg[512];
  sdegpu_last_error(msg, sizeof msg);
  printf("invalid problem: status %d, \"%s\"\n", rc, msg);
  if (rc != SDEGPU_EUNSUPPORTED || strlen(msg) == 0) return 8;
  printf("ok\n");
  return 0;
}
