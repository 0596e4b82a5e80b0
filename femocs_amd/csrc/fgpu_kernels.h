// Launch helpers of the gfx950 kernels (fgpu_kernels.hip).
#pragma once
#include "fgpu_internal.h"

namespace fgpu {
template <class T>
void launch_elem(hipStream_t s, const Desc *D, const ETab *tab, int nblk, bool coef, bool maskin, bool active_only,
                 int scal_slot, size_t lds, bool affine_all = false);
template <class T> void launch_elem_diag(hipStream_t s, const Desc *D, const ETab *tab, int nblk, bool coef, size_t lds);
void launch_elem_rhs(hipStream_t s, const Desc *D, const ETab *tab, int nblk, const double *fq, size_t lds);
void launch_elem_vol(hipStream_t s, const Desc *D, const ETab *tab, int nblk, size_t lds);
void launch_charge(hipStream_t s, int n, const double *b, const double *vol, double *rho);
template <class T> void launch_gather_cg(hipStream_t s, const Desc *D, const int2 *tab, int nblk);
template <class T> void launch_update_cg(hipStream_t s, const Desc *D, const int2 *tab, int nblk);
template <class T> void launch_gather_res(hipStream_t s, const Desc *D, const int2 *tab, int nblk);
template <class T> void launch_gather_plain(hipStream_t s, const Desc *D, const int2 *tab, int nblk, bool accum);
void launch_scal_alpha(hipStream_t s, const Desc *D, int n);
void launch_scal_beta(hipStream_t s, const Desc *D, int n, int reliable);
void launch_scal_res(hipStream_t s, const Desc *D, const Desc *D32, int n, int mixed);
void launch_scal_restart(hipStream_t s, const Desc *D32, const Desc *D64, int n, int first);
void launch_mixed_restart(hipStream_t s, const Desc *D32, const Desc *D64, const int2 *tab, int nblk, int first);
void launch_mixed_acc(hipStream_t s, const Desc *D32, const Desc *D64, const int2 *tab, int nblk);
void launch_collect(hipStream_t s, const Desc *D, int n, int slot, double *out);
void launch_coarse(hipStream_t s, const Desc *D32, const Desc *D64, const int2 *tab, int nblk, int mode, int max_nc);
void launch_coarse64(hipStream_t s, const Desc *D, const int2 *tab, int nblk, int max_nc);
void launch_gather_cheb(hipStream_t s, const Desc *D, const int2 *tab, int nblk, int j);
void launch_coarse_indicator(hipStream_t s, int n, const unsigned char *mask, const int *color, int shift, int c,
                             double *u);
void launch_coarse_sums(hipStream_t s, int n, const unsigned char *mask, const double *q, double *out);
int gj_block();
void launch_gj_block(hipStream_t s, int n, int k, int bb, const double *A, double *B, double *Pi, double *V,
                     double *W);
void launch_to_half(hipStream_t s, int n, int ld, const double *a, double scale, _Float16 *out);
void launch_enforce_dirichlet(hipStream_t s, int n, const unsigned char *mask, const double *g, double *x);
void launch_lift(hipStream_t s, int n, const unsigned char *mask, const double *g, const double *Ag, const double *diag,
                 double *b, double *x);
void launch_invdiag(hipStream_t s, int n, const unsigned char *mask, const double *diag, double *inv64, float *inv32);
void launch_set_const(hipStream_t s, int n, double v, double *x);
void launch_face_load(hipStream_t s, int nb, const int *bf_icell, const int *bf_face, const int *bf_id, int bid,
                      const double *fval, const int *cells, const double *X, double *fslot);
void launch_face_gather(hipStream_t s, int n, const int *fgptr, const int *fgidx, const double *fslot, double *b);
void launch_minmax(hipStream_t s, int n, const double *x, double *out);
int center_det_blocks(int n);
hipError_t upload_affine_table(hipStream_t s);
int metric_blocks(int nc);
void launch_metric(hipStream_t s, int nc, const int *cells, const unsigned char *layout, const double *X, float *metric,
                   double *devmax);
void launch_heat_terms(hipStream_t s, int nc, const int *cells, const double *X, const double *T, const double *phic,
                       double gamma, double lorentz, double theta, double *kq, double *gq, float *kq32, float *gq32,
                       double *f, double *kx);
void launch_sigma_terms(hipStream_t s, int nc, const int *cells, const double *T, double sigma0, double alpha, double *kq,
                        double *gq, float *kq32, float *gq32);
void launch_motion(hipStream_t s, int rows, const int *idof, const int *ptr, const int *ctrl, const double *w,
                   const double *base, const double *disp, double *X64, float *X32);
void launch_scatter_coords(hipStream_t s, int n, const int *idx, const double *xyz, double *X64, float *X32);
void launch_center_det_min(hipStream_t s, int n, const int *cells, const double *X, double *out);
void launch_min_reduce(hipStream_t s, int n, const double *in, double *out);
void launch_grad_export(hipStream_t s, int n, const int *v2c, const int *v2n, const int *cells, const double *X,
                        const double *u, double *grad);
}  // namespace fgpu
