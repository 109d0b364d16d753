/*
 * rgp.h - C ABI of the MI355X (gfx950) palette-weight Gaussian rasterizer.
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream, and
 * returns RGP_OK (0) or a negative error code; rgp_last_error() describes the
 * last failure of the calling thread.  No entry point allocates device memory:
 * buffers beyond the outputs come in through a caller-provided workspace
 * (rgp_bin_workspace_bytes).  All kernels are enqueued on the given stream and
 * no entry point synchronises the host, so sequences of calls can be captured
 * into a HIP graph.  Pointer arguments marked "nullable" may be NULL.
 *
 * Each declaration cites the reference interface it replaces
 * (tedchao/ReparamGS-Palette, gsplatcu/ext.cpp bindings implemented in
 * gsplatcu/gausplat.cu and gsplatcu/kernel.cu).
 */
#ifndef RGP_H_
#define RGP_H_

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RGP_OK 0
#define RGP_ERR_ARG (-1)    /* invalid argument (size, K, pointer) */
#define RGP_ERR_LAUNCH (-2) /* a kernel launch failed */

#define RGP_MAX_DIM 8 /* MAX_SPLAT_DIM (kernel.cu:12): channels per Gaussian */
#define RGP_MAX_VIEWS 16 /* cameras of one multi-view launch (the *_views entry points) */

/* Camera for the fused entry points, resident in device memory so that a
 * captured graph can be replayed with a new pose.  Fields follow the
 * reference Camera (gsplat/gausplat_dataset.py:21-33); tan_fovx/y are
 * width/(2 fx) and height/(2 fy) as gausplat.cu:357-358 computes them. */
typedef struct rgp_camera {
  float Rcw[9];
  float tcw[3];
  float twc[3];
  float fx, fy, cx, cy;
  float tan_fovx, tan_fovy;
  int32_t width, height;
  int32_t pad[9]; /* 128 bytes: arrays of cameras (the *_views entry points) have a 128-byte stride */
} rgp_camera;

const char* rgp_last_error(void);
int rgp_version(void);

/* ---- per-Gaussian geometry (palette path calls them with calc_J = False) -- */
/* The Jacobian outputs (calc_J = True) are nullable; when given they must be
 * zero-filled by the caller like the reference's torch::full(..., 0.0)
 * allocations (gausplat.cu:312-316, 350-354, 392-393, 442-446, 601-602):
 * entries the reference does not write (culled Gaussians, the structural
 * zeros of du/dpc) keep that fill. */

/* gsc.project, ext.cpp:94-101 / gausplat.cu:381-419 / kernel.cu:618-663.
 * pws (N,3); Rcw (3,3), tcw (3) device.  Outputs us (N,2), pcs (N,3),
 * depths (N) must be zero-filled by the caller (the reference allocates them
 * with torch::full 0); culled Gaussians get depth -1 and keep zeros.
 * du_dpcs (N,2,3), kernel.cu:656-660: entries 0, 2, 4, 5. */
int rgp_project(int n, const float* pws, const float* Rcw, const float* tcw, float fx, float fy,
                float cx, float cy, float* us, float* pcs, float* depths, float* du_dpcs,
                hipStream_t stream);

/* gsc.computeCov3D, ext.cpp:78-81 / gausplat.cu:298-331 / kernel.cu:422-501.
 * rots (N,4) (w,x,y,z), scales (N,3) linear -> cov3ds (N,6) (zero-filled).
 * dcov3d_drots (N,6,4), dcov3d_dscales (N,6,3), kernel.cu:463-500: both or
 * neither. */
int rgp_compute_cov3d(int n, const float* rots, const float* scales, const float* depths,
                      float* cov3ds, float* dcov3d_drots, float* dcov3d_dscales,
                      hipStream_t stream);

/* gsc.computeCov2D, ext.cpp:83-91 / gausplat.cu:333-379 / kernel.cu:508-611.
 * dcov2d_dcov3ds (N,3,6), dcov2d_dpcs (N,3,3), kernel.cu:557-607: both or
 * neither (the +-1.3 tan_fov clamp is not differentiated, as there). */
int rgp_compute_cov2d(int n, const float* cov3ds, const float* pcs, const float* Rcw,
                      const float* depths, float fx, float fy, float width, float height,
                      float* cov2ds, float* dcov2d_dcov3ds, float* dcov2d_dpcs,
                      hipStream_t stream);

/* gsc.inverseCov2D, ext.cpp:73-75 / gausplat.cu:590-617 / kernel.cu:379-415.
 * depths is updated in place (NaN determinant -> -1).  dcinv2d_dcov2ds
 * (N,3,3), kernel.cu:402-414, for the Gaussians that survive. */
int rgp_inverse_cov2d(int n, const float* cov2ds, float* depths, float* cinv2ds, int32_t* areas,
                      float* dcinv2d_dcov2ds, hipStream_t stream);

/* gsc.sh2Weight, ext.cpp:103-107 / gausplat.cu:425-462 / kernel.cu:670-870.
 * shs (N, sh_dim*K) band-major; twc (3) device.  basis (N, sh_dim) nullable
 * (the calc_J output dweight_dshs); dweight_dpws (N,K,3) nullable, the
 * reference's dW/dpw (kernel.cu:769-869), including its dY_9/dy term
 * C30 (3x^2 - 2y^2) (kernel.cu:840; the exact derivative is
 * C30 (3x^2 - 3y^2)). */
int rgp_sh2weight(int n, int K, int sh_dim, const float* shs, const float* pws, const float* twc,
                  float* weights, float* basis, float* dweight_dpws, hipStream_t stream);

/* gsc.sh2Weight_fast, ext.cpp:116-124 / gausplat.cu:469-509 / kernel.cu:884-987,
 * generalised from rank 1 to rank r: A (N,P,r), B (N,r,Q), P*Q = 15*K.
 * dweight_ddc (N,K) and ylm (N,15) nullable (calc_J outputs). */
int rgp_sh2weight_fast(int n, int K, int P, int Q, int rank, const float* dc, const float* A,
                       const float* B, const float* pws, const float* twc, float* weights,
                       float* dweight_ddc, float* ylm, hipStream_t stream);

/* gsc.sh2Weight_fast_B, ext.cpp:127-133 / gausplat.cu:512-541 /
 * kernel.cu:994-1032, rank r.  dA (N,P,r), dB (N,r,Q) fully overwritten. */
int rgp_sh2weight_fast_backward(int n, int K, int P, int Q, int rank, const float* A,
                                const float* B, const float* ylm, const float* dloss_dw,
                                float* dA, float* dB, hipStream_t stream);

/* gsc.sh2Color, ext.cpp:109-112 / gausplat.cu:547-584 / kernel.cu:1040-1169
 * (colour path, used to render frozen targets).  calc_J outputs (both or
 * neither): dcolor_dshs (N, sh_dim), dcolor_dpws (N, 3, 3). */
int rgp_sh2color(int n, int sh_dim, const float* shs, const float* pws, const float* twc,
                 float* colors, float* dcolor_dshs, float* dcolor_dpws, hipStream_t stream);

/* ---- binning: getRects + scan + createKeys + sort + getRanges ----------- */

/* Bytes of workspace for n Gaussians, a pair capacity and a tile grid. */
size_t rgp_bin_workspace_bytes(int n, int64_t capacity, int width, int height);

/* getRects (kernel.cu:127-162, gausplat.cu:189-197) into the workspace:
 * rects and per-Gaussian tile counts; culls in place (depths -1, areas 0). */
int rgp_bin_rects(int n, const float* us, int32_t* areas, float* depths, int width, int height,
                  void* ws, int64_t capacity, hipStream_t stream);

/* inclusive scan of the counts (gausplat.cu:199-201), and for grids up to
 * 32768 tiles the bucketing of the pairs that rgp_bin_sort completes.  The
 * pair count M and an overflow flag (M > capacity) are left in the
 * workspace header: rgp_bin_header() gives their device addresses.  Runs
 * after rgp_bin_rects or rgp_preprocess_weight(_views) on the same
 * workspace (they clear the header and write the counts and rects). */
int rgp_bin_scan(int n, int width, int height, void* ws, int64_t capacity, hipStream_t stream);

/* createKeys + stable (tile, uint32(depth*1000)) sort + getRanges
 * (kernel.cu:96-187, gausplat.cu:203-226).  ids_out (capacity) receives the
 * Gaussian id of every sorted pair, keys_out (capacity, nullable) the sorted
 * keys, ranges_out (tiles, 2) the [start, end) of every tile ((0,0) when the
 * tile is empty, and for every tile when M == 1 as in the reference).  When
 * M > capacity all ranges are (0,0) and the overflow flag is set. */
int rgp_bin_sort(int n, const float* depths, int width, int height, void* ws, int64_t capacity,
                 int32_t* ids_out, uint64_t* keys_out, int32_t* ranges_out, hipStream_t stream);

/* Device addresses of the header words: [0] M (u32), [1] overflow flag. */
const uint32_t* rgp_bin_header(const void* ws);

/* ---- splat records and compositing --------------------------------------- */

/* Floats per splat record for dim channels: 8 + 4*ceil(dim/4). */
int rgp_record_floats(int dim);

/* Pack us (N,2), cinv2ds (N,3), alphas (N), feat (N,dim) into records. */
int rgp_pack_records(int n, int dim, const float* us, const float* cinv2ds, const float* alphas,
                     const float* feat, float* records, hipStream_t stream);

/* draw_weight / draw_color (kernel.cu:194-372, gausplat.cu:89-100, 228-240):
 * per-tile front-to-back compositing of dim channels.  Writes every pixel:
 * layer (dim,H,W), contrib (H,W), final_tau (H,W); empty tiles get 0, 0 and
 * empty_tau (0 for splat_weight, gausplat.cu:178; 1 for splat_color, :39).
 * strips (nullable, uint8, one per list position like ids): for every entry
 * a tile composites, the mask of its four 16x4 row strips (bit j: rows
 * 4j..4j+3 of the tile) in which some pixel took a contribution; entries after
 * the batch of 64 in which the tile's pixels all stopped are not written.
 * rgp_draw_backward_replay reads it back.  No reference counterpart.
 * contrib (nullable): when null the contribution index is neither tracked
 * nor written (the replay backward with strips does not read it, and frames
 * that only show the images do not need it); layer and final_tau are the
 * same either way. */
int rgp_draw_forward(int width, int height, int dim, const int32_t* ranges, const int32_t* ids,
                     const float* records, float empty_tau, float* layer, int32_t* contrib,
                     float* final_tau, uint8_t* strips, hipStream_t stream);

/* rgp_draw_forward plus two counters added into counts (int64[2], caller
 * zero-fills): list entries evaluated per pixel before it stops, summed, and
 * contributing entries (alpha' >= 0.002), summed.  Instrumentation of the
 * forward's algorithmic work; no counterpart in the reference. */
int rgp_draw_forward_counts(int width, int height, int dim, const int32_t* ranges,
                            const int32_t* ids, const float* records, float empty_tau,
                            float* layer, int32_t* contrib, float* final_tau, int64_t* counts,
                            hipStream_t stream);

/* draw_weight_B (kernel.cu:1291-1379, gausplat.cu:252-292): reverse-order
 * backward of the channel compositing.  dfeat (N, dim) must be zero-filled;
 * gradients are added with wavefront reductions then one atomic per
 * (wavefront, Gaussian, channel). */
int rgp_draw_backward(int width, int height, int dim, const int32_t* ranges, const int32_t* ids,
                      const float* records, const int32_t* contrib, const float* final_tau,
                      const float* dlayer, float* dfeat, hipStream_t stream);

/* Colour backward with geometry gradients: draw_color_B (kernel.cu:1176-1289,
 * gausplat.cu:111-155) and the vanilla trainer's drawB
 * (vanilla_3dgs/gsplatcu/kernel.cu:811-971), which adds dL/dfinal_tau.
 * records are rgp_pack_records of (us, cinv2ds, alphas, colors) with dim 3;
 * dimage (3,H,W); dfinal_tau (H,W) or NULL.  dus (N,2), dcinv (N,3),
 * dalphas (N), dcolors (N,3) must be zero-filled; the derivatives of the 0.99
 * clamp and of max(0, .) are not taken, as in the reference. */
int rgp_draw_color_backward(int width, int height, const int32_t* ranges, const int32_t* ids,
                            const float* records, const int32_t* contrib, const float* final_tau,
                            const float* dimage, const float* dfinal_tau, float* dus, float* dcinv,
                            float* dalphas, float* dcolors, hipStream_t stream);

/* The vanilla training step's backward chain (vanilla_3dgs/gsplat/gsmodel.py:
 * 59-73) in one kernel instead of batched matrix products: from the colour
 * backward's dus (N,2), dcinv (N,3), dcolors (N,3) and the forward's calc_J
 * Jacobians (shapes of gausplat.cu's outputs), writes dpws (N,3),
 * dshs (N, sh_dim*3), drots (N,4), dscales (N,3).  Rcw is the (3,3) device
 * camera rotation; every product accumulates from 0 in index order. */
int rgp_vanilla_chain(int n, int sh_dim, const float* Rcw, const float* dus, const float* dcinv,
                      const float* dcolors, const float* du_dpcs, const float* dcinv_dcov2d,
                      const float* dcov2d_dcov3d, const float* dcov2d_dpc, const float* dcov3d_drot,
                      const float* dcov3d_dscale, const float* dcolor_dshs, const float* dcolor_dpw,
                      float* dpws, float* dshs, float* drots, float* dscales, hipStream_t stream);

/* Replay form of draw_weight_B's result: the forward recurrence is replayed
 * front to back and every contributing (pixel, entry) adds
 * dlayer[k,p] * tau * alpha' to dfeat, tau being the transmittance in front
 * of the entry (the factor the forward applied to the features).  Needs no
 * contrib / final_tau; the contributing set is the forward's.  The values
 * differ from rgp_draw_backward only by the rounding of the reference's
 * division-recovered tau (kernel.cu:1371).  dfeat zero-filled by caller.
 * strips (nullable): the masks rgp_draw_forward wrote for the same ranges,
 * ids and records; only those strips of each entry are evaluated (the same
 * values, fewer evaluations).  NULL: every strip the entry can reach. */
int rgp_draw_backward_replay(int width, int height, int dim, const int32_t* ranges,
                             const int32_t* ids, const float* records, const uint8_t* strips,
                             const float* dlayer, float* dfeat, hipStream_t stream);

/* ---- fused paths (no reference counterpart as single calls; they produce
 * the same values as the reference chains they replace) ------------------ */

/* project + computeCov3D + computeCov2D + sh2Weight(_fast) + inverseCov2D +
 * getRects of WeightGSFunctionTest/Fast.forward (gsmodel.py:107-118 /
 * 155-170) in one pass.  mode 0: full rank, shs (N, P*K) with P the number
 * of SH coefficients per channel (1, 4, 9 or 16; Q and rank unused); mode 1:
 * low rank, dc (N,K), A (N,P,r), B (N,r,Q); mode 2: full rank held as the
 * reference's two leaves, dc = low_shs_w (N,K) and shs = high_shs_w
 * (N,(P-1)*K), giving the values of mode 0 on their concatenation.  Writes depths (N, with -1 culls), records
 * (N, rgp_record_floats(K)), and rects/counts into ws.  us/cinv2ds/areas/
 * weights (nullable) receive the unfused outputs for inspection. */
int rgp_preprocess_weight(int n, int K, int mode, int P, int Q, int rank,
                          const rgp_camera* cam, const float* pws, const float* rots,
                          const float* scales, const float* alphas, const float* shs,
                          const float* dc, const float* A, const float* B, float* depths,
                          float* records, void* ws, int64_t capacity, int width, int height,
                          float* us, float* cinv2ds, int32_t* areas, float* weights,
                          hipStream_t stream);

/* ---- multi-view batches (no counterpart in the reference, which renders
 * one camera per call: gsmodel.py:104-200 called once per view).  `views`
 * cameras of one resolution are processed in one launch per stage, each
 * view producing exactly what the single-view entry points produce for its
 * camera.  Per-view arrays are contiguous: view v's depths at depths + v*n,
 * records at records + v*n*rgp_record_floats(K), workspace at
 * ws + v*ws_stride (ws_stride a multiple of 256 and at least
 * rgp_bin_workspace_bytes), ids at ids + v*capacity, ranges at
 * ranges + v*tiles*2, images at layer + v*dim*H*W, contrib / final_tau at
 * + v*H*W, gradients at dfeat + v*n*dim. -------------------------------- */

/* rgp_preprocess_weight for `views` cameras (cams: device array) in one
 * pass over the Gaussians: the per-Gaussian parameters are read from HBM
 * once for all views. */
int rgp_preprocess_weight_views(int views, int n, int K, int mode, int P, int Q, int rank,
                                const rgp_camera* cams, const float* pws, const float* rots,
                                const float* scales, const float* alphas, const float* shs,
                                const float* dc, const float* A, const float* B, float* depths,
                                float* records, void* ws, size_t ws_stride, int64_t capacity, int width,
                                int height, hipStream_t stream);

/* rgp_bin_scan / rgp_bin_sort of every view (one launch sequence per view). */
int rgp_bin_scan_views(int views, int n, int width, int height, void* ws, size_t ws_stride, int64_t capacity,
                       hipStream_t stream);
int rgp_bin_sort_views(int views, int n, const float* depths, int width, int height, void* ws,
                       size_t ws_stride, int64_t capacity, int32_t* ids, int32_t* ranges, hipStream_t stream);

/* rgp_draw_forward of every view in one launch (views x tiles blocks);
 * id_stride = capacity, n the Gaussian count (record stride); strips
 * (nullable) has the stride of ids; contrib is nullable as for
 * rgp_draw_forward. */
int rgp_draw_forward_views(int views, int width, int height, int dim, const int32_t* ranges,
                           const int32_t* ids, int64_t id_stride, const float* records, int64_t n,
                           float empty_tau, float* layer, int32_t* contrib, float* final_tau,
                           uint8_t* strips, hipStream_t stream);

/* rgp_draw_backward_replay of every view in one launch; view v adds into its
 * own rows dfeat + v*n*dim (zero-filled by the caller); strips (nullable) as
 * rgp_draw_forward_views wrote them. */
int rgp_draw_backward_replay_views(int views, int width, int height, int dim, const int32_t* ranges,
                                   const int32_t* ids, int64_t id_stride, const float* records, int64_t n,
                                   const uint8_t* strips, const float* dlayer, float* dfeat,
                                   hipStream_t stream);

/* The SH chains of the backward for a multi-view batch: the gradient of the
 * coefficients summed over the views, each view with its own basis
 * Y(pw, twc_v) and weight gradient dw + v*n*K (cams: device array).
 * Full rank (gsmodel.py:140-143) and low rank (gsmodel.py:194-198). */
int rgp_sh_backward_full_views(int views, int n, int K, int sh_dim, const rgp_camera* cams, const float* pws,
                               const float* dw, float* dshs, hipStream_t stream);
int rgp_sh_backward_lowrank_views(int views, int n, int K, int P, int Q, int rank, const rgp_camera* cams,
                                  const float* pws, const float* A, const float* B, const float* dw, float* ddc,
                                  float* dA, float* dB, hipStream_t stream);

/* Full-rank SH backward of gsmodel.py:140-143: dshs[n, lm*K+k] =
 * dw[n,k] * Y_lm(n), Y recomputed from pws and twc (device camera). */
int rgp_sh_backward_full(int n, int K, int sh_dim, const rgp_camera* cam, const float* pws,
                         const float* dw, float* dshs, hipStream_t stream);

/* No counterpart in the reference: the SH backward above (gsmodel.py:140-143)
 * fused with the update torch.optim.Adam (no amsgrad) applies to the full-rank
 * coefficients in the finetune step (finetune.py:346-399), so dL/dshs is never
 * stored.  shs, exp_avg and exp_avg_sq are (N, sh_dim*K) float32, 16-byte
 * aligned, updated in place; step is the Adam step count after this update
 * (>= 1); 1 - beta, the bias corrections and lr / bc1 are formed in double
 * from the double hyper-parameters, as torch does, then rounded to float. */
int rgp_sh_backward_adam_full(int n, int K, int sh_dim, const rgp_camera* cam, const float* pws,
                              const float* dw, float* shs, float* exp_avg, float* exp_avg_sq,
                              int64_t step, double lr, double beta1, double beta2, double eps,
                              double weight_decay, hipStream_t stream);

/* No counterpart in the reference: the same fused update for the
 * reference's two coefficient leaves (gsplat/utils.py:157-158 concatenates
 * low_shs_w (N,K) and high_shs_w (N,15K) before every render): low holds
 * band 0 and high bands 1-15, each with its own Adam moments; the values are
 * those of rgp_sh_backward_adam_full on the concatenated tensor. */
int rgp_sh_backward_adam_split(int n, int K, const rgp_camera* cam, const float* pws, const float* dw, float* low,
                               float* low_exp_avg, float* low_exp_avg_sq, float* high, float* high_exp_avg,
                               float* high_exp_avg_sq, int64_t step, double lr, double beta1, double beta2,
                               double eps, double weight_decay, hipStream_t stream);

/* The low-rank SH backward fused with torch.optim.Adam's update of the three
 * factor tensors, for WeightGSFunctionFast's register path (Q = 15, P = K,
 * rank 1 or 2): replaces rgp_sh_backward_lowrank (gsmodel.py:194-198 and
 * sh2Weight_fast_B, kernel.cu:991-1032) followed by torch.optim.Adam on
 * shs_dc (N,K), shs_A (N,K,rank) and shs_B (N,rank,15) (the optimizer of
 * finetune.py:303).  Each tensor has its own exp_avg / exp_avg_sq;
 * the update is torch.optim.Adam's, bit-identical to rgp_sh_backward_lowrank
 * then rgp_adam on each tensor.  All nine arrays 16-byte aligned; other
 * shapes return RGP_ERR_ARG (use the unfused pair). */
int rgp_sh_backward_adam_lowrank(int n, int K, int P, int Q, int rank, const rgp_camera* cam, const float* pws,
                                 const float* dw, float* dc, float* dc_exp_avg, float* dc_exp_avg_sq, float* A,
                                 float* A_exp_avg, float* A_exp_avg_sq, float* B, float* B_exp_avg,
                                 float* B_exp_avg_sq, int64_t step, double lr, double beta1, double beta2,
                                 double eps, double weight_decay, hipStream_t stream);

/* rgp_sh_backward_adam_lowrank for a multi-view batch: the factor gradients
 * are summed over `views` cameras (cams: device array, dw: (views, N, K)) as
 * rgp_sh_backward_lowrank_views forms them, then updated as above; equal bit
 * for bit to rgp_sh_backward_lowrank_views followed by rgp_adam per tensor.
 * No reference counterpart (the reference renders one camera per call). */
int rgp_sh_backward_adam_lowrank_views(int views, int n, int K, int P, int Q, int rank, const rgp_camera* cams,
                                       const float* pws, const float* dw, float* dc, float* dc_exp_avg,
                                       float* dc_exp_avg_sq, float* A, float* A_exp_avg, float* A_exp_avg_sq,
                                       float* B, float* B_exp_avg, float* B_exp_avg_sq, int64_t step, double lr,
                                       double beta1, double beta2, double eps, double weight_decay,
                                       hipStream_t stream);

/* No counterpart in the reference either: one torch.optim.Adam step (the
 * optimizer of finetune.py:346-399) over a flat float32 tensor of n elements
 * with its gradient, in one pass; the same element update and scalars as
 * rgp_sh_backward_adam_full. */
int rgp_adam(int64_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t step,
             double lr, double beta1, double beta2, double eps, double weight_decay, hipStream_t stream);

/* Low-rank SH backward of gsmodel.py:194-198 (dL/ddc = dw * SH_C0 and
 * sh2Weight_fast_B), Y recomputed from pws and twc. */
int rgp_sh_backward_lowrank(int n, int K, int P, int Q, int rank, const rgp_camera* cam,
                            const float* pws, const float* A, const float* B, const float* dw,
                            float* ddc, float* dA, float* dB, hipStream_t stream);

/* ---- palette mix (no reference kernel; equals bake_palette_edit +
 * render_vanilla, gsmodel.py:488-509 + 13-26, by linearity) -------------- */

/* rgb (3,H,W) = sum_k palette[k] * W[k] for a batch of `count` images:
 * W pointers, rgb pointers, pixel counts and palettes (count, K, 3) in
 * device arrays.  One launch serves every image of the batch. */
int rgp_mix_batched(int count, int K, const float* const* W, float* const* rgb,
                    const int64_t* hw, const float* palettes, int64_t max_hw, hipStream_t stream);

/* Single image convenience form. palette (K,3) device. */
int rgp_mix(int K, int64_t hw, const float* W, const float* palette, float* rgb,
            hipStream_t stream);

/* Backward of the mix: dW[k,p] = sum_c palette[k,c] drgb[c,p];
 * dpalette[k,c] += sum_p W[k,p] drgb[c,p] (dpalette zero-filled by caller). */
int rgp_mix_backward(int K, int64_t hw, const float* W, const float* palette, const float* drgb,
                     float* dW, float* dpalette, hipStream_t stream);

/* ---- image loss of the finetune-step workload: (1 - lam) L1 + lam (1 - SSIM)
 * with the 11x11 Gaussian window of gsplat/pytorch_ssim.py:17-66 and
 * vanilla_3dgs/gsplat/pytorch_ssim.py:64-67 (no gsplatcu counterpart; the
 * reference computes it with torch conv2d). ------------------------------ */

/* Length (float64 elements) of the sums workspace of rgp_l1_ssim_forward:
 * one pair per 32x32 output tile and channel. */
int64_t rgp_l1_ssim_sums(int channels, int height, int width);

/* img, target (C,H,W); window11 a host array of the 11 normalised taps.
 * abc (3,C,H,W) receives the per-pixel SSIM partials the backward needs;
 * sums (float64[rgp_l1_ssim_sums(C,H,W)], 16-byte aligned, no initialisation needed) the
 * per-tile partial sums of the SSIM map and of |img - target|, added in a
 * fixed order (the loss value is deterministic);
 * loss (device float) the loss value. */
int rgp_l1_ssim_forward(int channels, int height, int width, const float* img,
                        const float* target, const float* window11, float C1, float C2, float lam,
                        float* abc, double* sums, float* loss, hipStream_t stream);

/* dimg (C,H,W) = grad_loss (device float) * dloss/dimg. */
int rgp_l1_ssim_backward(int channels, int height, int width, const float* img,
                         const float* target, const float* abc, const float* window11, float lam,
                         const float* grad_loss, float* dimg, hipStream_t stream);

/* ---- barycentric palette loss of the finetune step (no gsplatcu
 * counterpart: the reference builds it from torch operations,
 * barycentric_coords_3d at gsplat/gsmodel.py:219-285 and the loss terms at
 * finetune.py:353-392).  K_c chromatic palette colours (1 <= Kc <= 6), the
 * weight images have Kc + 2 channels (black, white, the chromatic ones). -- */

/* Float64 accumulators the two entry points below take as workspace. */
size_t rgp_bary_workspace_bytes(int Kc);

/* image (3,H,W) the frozen render C; w_tilde (Kc+2,H,W) the rendered weights;
 * gt_alpha (H,W) or NULL (with it: C / max(a, 1e-3) clamped to [0,1],
 * W_bary * a, and lambda_alpha * mean((sum_k W_tilde - a)^2)); palette (Kc,3);
 * clamp != 0 applies W_bary <- max(W_bary,0) / max(sum, 1e-8).
 * Loss L = mean((W_tilde - W_bary)^2) + lambda_pos mean(relu(-W_tilde))
 * [+ lambda_alpha L_alpha].  Writes dw_tilde = dL/dW_tilde (Kc+2,H,W),
 * dpalette = dL/dpalette (Kc,3), loss[4] = (L, L_geo, L_pos, L_alpha) and,
 * if w_bary is not NULL, the target weights W_bary (Kc+2,H,W). */
int rgp_bary_loss(int width, int height, int Kc, const float* image, const float* w_tilde,
                  const float* gt_alpha, const float* palette, float tau, int clamp,
                  float lambda_pos, float lambda_alpha, float* w_bary, float* dw_tilde,
                  float* dpalette, float* loss, void* workspace, hipStream_t stream);

/* rgp_bary_loss without the workspace argument.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The float64 accumulators live in a library-owned device array of
 * (RGP_MAX_DIM - 2) * 16 + 3 doubles (792 bytes) per GPU, held until the
 * process ends; its address is looked up on the first call on each device
 * (make that first call outside a graph capture).  Kc above RGP_MAX_DIM - 2
 * is refused.  The stream must belong to the current device.  Every call on
 * one GPU shares that GPU's accumulators (cleared by a fill kernel at the
 * start of each call), so calls that overlap on one GPU race.  Allocates
 * nothing. */
int rgp_bary_loss_noalloc(int width, int height, int Kc, const float* image, const float* w_tilde,
                          const float* gt_alpha, const float* palette, float tau, int clamp,
                          float lambda_pos, float lambda_alpha, float* w_bary, float* dw_tilde,
                          float* dpalette, float* loss, hipStream_t stream);

/* barycentric_coords_3d(C, P, tau) alone: w_bary (Kc+2,H,W). */
int rgp_bary_coords(int width, int height, int Kc, const float* image, const float* palette,
                    float tau, float* w_bary, hipStream_t stream);

/* Its backward: dpalette (Kc,3) = sum over pixels of dw_bary . dW_bary/dP. */
int rgp_bary_coords_backward(int width, int height, int Kc, const float* image,
                             const float* palette, float tau, const float* dw_bary,
                             float* dpalette, void* workspace, hipStream_t stream);

/* Fill n 32-bit words with a value (graph-capture safe, no memset node). */
int rgp_fill_u32(void* dst, uint32_t value, int64_t n, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* RGP_H_ */
