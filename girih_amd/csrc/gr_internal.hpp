// gr_internal.hpp -- types shared by the host-side translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace gr {

struct Geom;

// One problem's geometry and device fields (see gr_runtime.hip for the layout).
struct Problem {
  int kernel = 0, dtype = 0, flags = 0;
  int nx = 0, ny = 0, nz = 0, R = 0, ncoef = 0;
  int xoff = 0, pitch = 0, rows = 0, planes = 0;
  int cxoff = 0;  // x offset of the roc2 / coefficient storage (xoff, or its own for bf16 storage)
  long long field_elems = 0;  // elements of one field allocation
  double c[5] = {0, 0, 0, 0, 0};
  double roc2s = 0;
  int strm_mode = -1;    // non-temporal sweep instantiations: -1 by working-set size, 0 never, 1 always
  void *roc2 = nullptr;  // device roc2 field base (x offset applied)
  void *coef = nullptr;  // device coefficient arrays base (x offset applied)
};

struct LaunchCfg {
  int zchunk = 0;   // z planes per workgroup (0 = whole box)
  int t_block = 1;  // fused time steps per pass
  int variant = 0;  // sweep tile variant (0 = default tile)
  int xcd_order = 1;  // renumber workgroups per XCD
};

int num_tile_variants(int kid, int dtype);
void tile_of_variant(int kid, int dtype, int variant, int *tx, int *ty);

size_t elem_size(int dtype);
// Whether a sweep of P should use non-temporal hints: by default, when its fields and
// coefficient arrays (the data one step streams) exceed what the caches keep between
// steps; P.strm_mode 0 / 1 forces the choice.
int streaming_sweep(const Problem &P);
int set_error(int code, const char *msg);
int set_errorf(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// One sweep u <- f(v); kernel 0 reads u^(n-1) from uold when given (else from u).
hipError_t sweep_fields(const Problem &P, void *u, const void *v, const void *src, long long src_idx,
                        const LaunchCfg &lc, hipStream_t st, const void *uold = nullptr);
// While set (only inside fused_tile), the fused-pass launchers store their output tile
// (x, y cells) here and return without launching.
extern thread_local int *g_fused_tile_probe;
// Output tile of the fused pass of depth tl that tblock_pass would launch for P and lc;
// false where there is no such pass.
bool fused_tile(const Problem &P, int tl, const LaunchCfg &lc, int *tx, int *ty);
hipError_t tblock_pass(const Problem &P, int tl, const void *X, const void *Y, void *ZO, void *ZN, const void *src,
                       long long src_idx, const LaunchCfg &lc, hipStream_t st);
int tb_depths(int kid, int dtype, int *out, int max);
// tl kernel-1 steps per pass: Y -> Z.
hipError_t jacobi_tb_pass(const Problem &P, int tl, const void *Y, void *Z, const void *src, long long src_idx,
                          const LaunchCfg &lc, hipStream_t st);
// Two kernel-0 steps per pass: (X, Y) = (u^(n-1), u^n) -> (ZO, ZN) = (u^(n+1), u^(n+2)).
hipError_t wave2_pass(const Problem &P, const void *X, const void *Y, void *ZO, void *ZN, const void *src,
                      long long src_idx, const LaunchCfg &lc, hipStream_t st);
hipError_t sweep_box(const Problem &P, const Geom &g, void *u, const void *v, const LaunchCfg &lc, hipStream_t st);

}  // namespace gr

// Handle internals for the batched stepper (gr_batch.hip).
struct gr_handle;
const gr::Problem &gr_internal_problem(const gr_handle *h);
void *gr_internal_field(const gr_handle *h, int logical);  // 0 U1, 1 U2 (x offset applied)
void *gr_internal_spare(gr_handle *h, int which);          // spare field 0/1, allocated on demand (or null)
int gr_internal_parity(const gr_handle *h);
long long gr_internal_napplied(const gr_handle *h);
const void *gr_internal_source(const gr_handle *h, int *n, long long *idx);
hipStream_t gr_internal_stream(const gr_handle *h);
const gr::LaunchCfg &gr_internal_launch_cfg(const gr_handle *h);
void gr_internal_advance(gr_handle *h, int napply, double ms);
// Two-step kernel-0 passes for the batch: prepare (spares allocated and given the
// fields' halo; *ok false when the fields' halos differ and passes may not run), the
// buffers {read as u^(n-1), read as u^n, written as u^(n+1), written as u^(n+2)} of
// pass number npasses of a call, and the state after npasses passes plus
// napply - 2*npasses single steps.
int gr_internal_fused_prepare(gr_handle *h, bool *ok);
void gr_internal_fused_buffers(const gr_handle *h, int npasses, void *buf[4]);
void gr_internal_fused_advance(gr_handle *h, int npasses, int napply, double ms);
