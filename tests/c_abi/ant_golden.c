/*
 * A plain C client of libpippin_hip.so, with no torch in the process: it binds
 * ROCm's own HIP runtime, as a Rust FFI crate would. It loads an MJCF file
 * through the library's loader, steps it through the C ABI and writes what it
 * saw, for the caller to compare with the golden Ant rollout.
 *
 *   ant_golden MODEL.xml IN.bin OUT.bin
 *
 * IN.bin : int64 n, T, nq, nu; double qpos0[n][nq]; double ctrl[T][n][nu]
 * OUT.bin: double qpos[T/10][n][nq] (after steps 10, 20, ...); int32 ncon[T][n]
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pippin_hip.h"

static int check(int rc, const char* what) {
  if (rc != PH_OK) {
    fprintf(stderr, "ant_golden: %s failed (%d): %s\n", what, rc, pippin_hip_last_error());
    exit(2);
  }
  return rc;
}

static void* read_exact(FILE* f, size_t bytes) {
  void* p = malloc(bytes ? bytes : 1);
  if (!p || fread(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "ant_golden: short input\n");
    exit(2);
  }
  return p;
}

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: ant_golden MODEL.xml IN.bin OUT.bin\n");
    return 2;
  }
  FILE* in = fopen(argv[2], "rb");
  if (!in) return 2;
  int64_t hdr[4];
  if (fread(hdr, sizeof hdr, 1, in) != 1) return 2;
  const int64_t n = hdr[0], T = hdr[1], nq = hdr[2], nu = hdr[3];
  if (n <= 0 || T <= 0 || T % 10 || nq <= 0 || nu < 0 || n > 4096 || T > 100000) return 2;
  double* qpos0 = read_exact(in, (size_t)(n * nq) * sizeof(double));
  double* ctrl = read_exact(in, (size_t)(T * n * nu) * sizeof(double));
  fclose(in);

  ph_model* model = NULL;
  check(pippin_hip_model_load_file(argv[1], &model), "model_load_file");
  ph_options opt;
  memset(&opt, 0, sizeof opt);
  opt.solver_override = -1;
  ph_sim* sim = NULL;
  check(pippin_hip_create(pippin_hip_model_desc(model), n, &opt, &sim), "create");
  pippin_hip_model_free(model);
  ph_dims dims;
  check(pippin_hip_dims(sim, &dims), "dims");
  if (dims.nq != nq || dims.nu != nu) {
    fprintf(stderr, "ant_golden: model has nq %lld nu %lld, input %lld %lld\n", (long long)dims.nq,
            (long long)dims.nu, (long long)nq, (long long)nu);
    return 2;
  }
  check(pippin_hip_set(sim, PH_FIELD_QPOS, 0, n, qpos0, n * nq), "set qpos");
  double* qout = calloc((size_t)(T / 10 * n * nq), sizeof(double));
  int32_t* ncon = calloc((size_t)(T * n), sizeof(int32_t));
  for (int64_t t = 0; t < T; t++) {
    check(pippin_hip_set(sim, PH_FIELD_CTRL, 0, n, ctrl + t * n * nu, n * nu), "set ctrl");
    check(pippin_hip_step(sim, 0, n, 1), "step");
    check(pippin_hip_ncon(sim, 0, n, ncon + t * n, n), "ncon");
    if (t % 10 == 9) check(pippin_hip_get(sim, PH_FIELD_QPOS, 0, n, qout + (t / 10) * n * nq, n * nq), "get qpos");
  }
  pippin_hip_destroy(sim);
  FILE* out = fopen(argv[3], "wb");
  if (!out) return 2;
  fwrite(qout, sizeof(double), (size_t)(T / 10 * n * nq), out);
  fwrite(ncon, sizeof(int32_t), (size_t)(T * n), out);
  fclose(out);
  printf("ant_golden: %lld envs x %lld steps through the C ABI\n", (long long)n, (long long)T);
  free(qpos0);
  free(ctrl);
  free(qout);
  free(ncon);
  return 0;
}
