/*
 * cpu_baseline_mpi.c -- MPI driver of the CPU baseline: the oracle's
 * restatement of the reference recipe timed the way the reference runs,
 * `mpiexec -n P`, one single-threaded rank per core.
 *
 * TEST AND BENCHMARK INFRASTRUCTURE ONLY: run by the CPU-baseline leg of
 * bench.py as the reported CPU path.  The product library never uses it.
 *
 * Each rank takes a contiguous block of the cells (the reference splits the
 * coarse cells over ranks, include/base/diffusion_problem_ms.tpp:52) and
 * builds them one after another with DiffusionProblemBasis<dim>::run()'s
 * recipe (orc_cell_build / orc_cell_build_general with CG + SSOR(omega),
 * absolute tolerance, iteration cap: diffusion_problem_basis.tpp:295-317,
 * :440-474), keeping no cell's fine system after its element matrix (the
 * reference keeps all of them, which does not fit the larger batches).  The
 * timed region is MPI_Barrier + MPI_Wtime around the loop, max over ranks
 * (the "basis initialization and computation" scope,
 * diffusion_problem_ms.tpp:44-45).
 *
 * Usage: cpu_baseline_mpi <input file>
 * Input (native byte order): int32 dim, n_cells, coef_kind, rhs_kind, general,
 * solver_max_iter; double rhs_value, omega, abs_tol; int32 n_fine[n_cells];
 * double geometry[n_cells][...] (box [2 dim], or the 2^dim vertices [2^dim][dim]
 * of a general cell); double coefficient tiles, concatenated in cell order.
 * Output (rank 0, stdout): one JSON line {"seconds", "cells", "ranks", "fails",
 * "iters_mean", "iters_max"}.
 */
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int orc_coef_stride(int dim, int kind);
int orc_n_nodes(int dim, int nf);
int orc_cell_build(int dim, int nf, const double *box, int coef_kind, const double *tile, int rhs_kind,
                   double rhs_value, int solver, double omega, double abs_tol, int max_iter, double *phi,
                   double *Ae, double *Fe, int *iters);
int orc_cell_build_general(int dim, int nf, const double *verts, int coef_kind, const double *tile, int rhs_kind,
                           double rhs_value, int solver, double omega, double abs_tol, int max_iter, double *phi,
                           double *Ae, double *Fe, int *iters);

static int read_all(FILE *f, void *p, size_t size, size_t n) { return fread(p, size, n, f) == n ? 0 : 1; }

int main(int argc, char **argv)
{
  MPI_Init(&argc, &argv);
  int rank = 0, ranks = 1;
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &ranks);
  FILE *f = argc == 2 ? fopen(argv[1], "rb") : NULL;
  int32_t hdr[6];
  double par[3];
  int bad = !f || read_all(f, hdr, sizeof(int32_t), 6) || read_all(f, par, sizeof(double), 3);
  const int dim = bad ? 2 : hdr[0], n = bad ? 0 : hdr[1], kind = hdr[2], rkind = hdr[3], general = hdr[4];
  const int max_iter = hdr[5];
  bad = bad || (dim != 2 && dim != 3) || n < 0;
  int32_t *nf = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  const int nb = 1 << dim, gsz = general ? nb * dim : 2 * dim;
  double *geo = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * gsz);
  size_t *off = (size_t *)malloc(sizeof(size_t) * ((size_t)n + 1));
  bad = bad || read_all(f, nf, sizeof(int32_t), (size_t)n) || read_all(f, geo, sizeof(double), (size_t)n * gsz);
  off[0] = 0;
  for (int c = 0; c < n && !bad; ++c)
    {
      size_t cells = 1;
      for (int k = 0; k < dim; ++k)
        cells *= (size_t)nf[c];
      off[c + 1] = off[c] + cells * (size_t)orc_coef_stride(dim, kind);
    }
  double *coef = (double *)malloc(sizeof(double) * (bad ? 1 : (off[n] > 0 ? off[n] : 1)));
  bad = bad || read_all(f, coef, sizeof(double), off[n]);
  if (f)
    fclose(f);
  if (bad)
    {
      if (rank == 0)
        fprintf(stderr, "cpu_baseline_mpi: cannot read the input file\n");
      MPI_Abort(MPI_COMM_WORLD, 2);
    }
  /* contiguous block of cells per rank */
  const int c0 = (int)((long long)n * rank / ranks), c1 = (int)((long long)n * (rank + 1) / ranks);
  double Ae[64], Fe[8];
  int it[8];
  long long its = 0, itmax = 0, fails = 0;
  MPI_Barrier(MPI_COMM_WORLD);
  const double t0 = MPI_Wtime();
  for (int c = c0; c < c1; ++c)
    {
      const int rc = general ? orc_cell_build_general(dim, nf[c], geo + (size_t)c * gsz, kind, coef + off[c], rkind,
                                                      par[0], 1, par[1], par[2], max_iter, NULL, Ae, Fe, it)
                             : orc_cell_build(dim, nf[c], geo + (size_t)c * gsz, kind, coef + off[c], rkind, par[0], 1,
                                              par[1], par[2], max_iter, NULL, Ae, Fe, it);
      fails += rc != 0;
      for (int i = 0; i < nb; ++i)
        {
          its += it[i];
          if (it[i] > itmax)
            itmax = it[i];
        }
    }
  const double dt = MPI_Wtime() - t0;
  double tmax = 0;
  long long sums[2] = {its, fails}, tot[2] = {0, 0}, imax = 0;
  MPI_Reduce(&dt, &tmax, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
  MPI_Reduce(sums, tot, 2, MPI_LONG_LONG, MPI_SUM, 0, MPI_COMM_WORLD);
  MPI_Reduce(&itmax, &imax, 1, MPI_LONG_LONG, MPI_MAX, 0, MPI_COMM_WORLD);
  if (rank == 0)
    printf("{\"seconds\": %.6f, \"cells\": %d, \"ranks\": %d, \"fails\": %lld, \"iters_mean\": %.3f, "
           "\"iters_max\": %lld}\n",
           tmax, n, ranks, tot[1], n > 0 ? (double)tot[0] / ((double)n * nb) : 0.0, imax);
  free(nf);
  free(geo);
  free(off);
  free(coef);
  MPI_Finalize();
  return 0;
}
