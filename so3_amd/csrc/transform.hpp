// transform.hpp -- host orchestration of one Wigner transform on the GPU.
#pragma once
#include <cstddef>

#include "kernels.hpp"
#include "so3/so3_types.h"

namespace so3g {

enum Method { kViaSsht = 0, kDirect = 1 };

// Throws HipError on invalid parameters or device failures.
void validate(const so3_parameters_t *p, Method method);
void run_inverse(const so3_parameters_t *p, Method method, const void *flmn_dev, void *f_dev,
                 hipStream_t st);
void run_forward(const so3_parameters_t *p, Method method, const void *f_dev, void *flmn_dev,
                 hipStream_t st);
void run_adjoint_inverse(const so3_parameters_t *p, const void *f_dev, void *flmn_dev, hipStream_t st);
void run_adjoint_forward(const so3_parameters_t *p, const void *flmn_dev, void *f_dev, hipStream_t st);
void prepare(const so3_parameters_t *p, Method method);
// The inverse on the library's static workspace (no allocation of any
// kind): needs a prepared plan and a resident Delta table, and a workspace
// of at most scratch_limit_bytes().
void run_inverse_noalloc(const so3_parameters_t *p, Method method, const void *flmn, void *f, hipStream_t st);
size_t scratch_limit_bytes();
// Autotuner (transform.cpp): times the tuned choices for these parameters on
// `st` and keeps the fastest; out = (inverse order, forward order, Pw),
// ms_out = (inverse contraction, forward contraction, forward beta) times.
void tune(const so3_parameters_t *p, Method method, hipStream_t st, int *out, double *ms_out);
void tune_get(const so3_parameters_t *p, Method method, int *out);
void tune_save(const char *path);
// MATLAB sample layout (layout.hip): to_mex f -> (re, im), else (re, im) -> f;
// im is null for real signals (and may be null for a real-valued complex input)
void launch_mex_layout(bool to_mex, void *f, double *re, double *im, const so3_parameters_t *p, hipStream_t st);
void tune_clear();
int tune_load(const char *path);

// Batched execution: plan, tables and the workspace size (double2 elements)
// of one transform (builds what is missing, so call it outside graph
// capture), then the stage sequences on a caller-provided workspace of at
// least that size.  The enqueue functions allocate nothing and expect
// validated parameters whose plan exists.
size_t workspace_elems(const so3_parameters_t *p, Method method);
void enqueue_inverse(const so3_parameters_t *p, Method method, const void *flmn_dev, void *f_dev, void *ws,
                     hipStream_t st);
void enqueue_forward(const so3_parameters_t *p, Method method, const void *f_dev, void *flmn_dev, void *ws,
                     hipStream_t st);
size_t workspace_bytes(const so3_parameters_t *p, Method method);
// Grouped batches: one transform split at its contraction.  contract_item
// fills the item of a grouped contraction launch (inverse: flmn `in` -> X1
// at ws; forward: G' at ws -> flmn `out`); enqueue_inverse_post runs the
// inverse's stages after its contraction, enqueue_forward_pre the forward's
// stages before it.  Same kernels and arithmetic as enqueue_inverse /
// enqueue_forward, so the results are bit-identical.
void contract_item(const so3_parameters_t *p, Method method, bool inverse, const void *in, void *out, void *ws,
                   CItem *it);
void enqueue_inverse_post(const so3_parameters_t *p, Method method, void *f, void *ws, hipStream_t st);
void enqueue_forward_pre(const so3_parameters_t *p, Method method, const void *f, void *flmn, void *ws,
                         hipStream_t st);
void release_all();

// Stage profiler: when enabled, every stage launch is bracketed by HIP
// events on its stream; stage_times() resolves them (synchronising) and
// returns accumulated milliseconds and launch counts per stage name.
void set_stage_timing(bool on);
int stage_times(const char **names, double *ms, int *count, int cap);
void reset_stage_times();

// true when every stored flmn entry is written by run_forward
bool forward_writes_all(const so3_parameters_t *p, Method method);

}  // namespace so3g
