/*
 * sbm_hip.h -- C ABI of libsbm_hip.so, the MI355X (gfx950) implementation of line2Dup's
 * matching path (meiqua/shape_based_matching).
 *
 * Every entry point takes plain pointers and sizes and returns an int status (SBM_OK == 0, or a
 * negative SBM_E* code; sbm_last_error() gives the message, per calling thread).  The comment on each
 * declaration names the reference interface it replaces (line2Dup.h / line2Dup.cpp / test.cpp).
 *
 * Images are 8-bit, 1 (grayscale) or 3 (BGR interleaved) channels, row-major with a row stride in
 * bytes.  The caller owns every buffer it passes; the library copies what it needs.  The library
 * owns the template bank (host copy plus a packed copy in HBM) and the match lists until freed.
 * One detector runs on one HIP stream on the device that was current when it first touched the
 * GPU; calls on one detector must not overlap (use one detector per thread, as with any stateful
 * object).  Nothing here falls back to the CPU: the matching entry points fail with SBM_ENODEV
 * when no GPU is visible.
 */
#ifndef SBM_HIP_H
#define SBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

/* status codes */
#define SBM_OK 0
#define SBM_EINVAL (-1)      /* bad argument */
#define SBM_ESIZE (-2)       /* image size not divisible at some level (line2Dup.cpp:612, 681, 724-725) */
#define SBM_EFEATURES (-3)   /* a template level holds >= 8192 features (line2Dup.cpp:1166, 1231);
                                >= 16384 with 16 orientations */
#define SBM_EDUPCLASS (-4)   /* readClass of a class already present (line2Dup.cpp:1424) */
#define SBM_EHIP (-5)        /* HIP runtime error */
#define SBM_EBUDGET (-6)     /* device-memory budget exceeded */
#define SBM_EIO (-7)         /* file / parse error */
#define SBM_ENOTFOUND (-8)   /* class or template not found (line2Dup.cpp:1366-1367) */
#define SBM_ETOOFEW (-9)     /* reserved */
#define SBM_ENODEV (-10)     /* no GPU visible */

/* match modes */
#define SBM_MODE_PLAIN 0     /* one coarse launch per template, as matchClass iterates them */
#define SBM_MODE_BATCHED 1   /* the whole bank in one persistent launch */

typedef struct sbm_detector sbm_detector;
typedef struct sbm_matches sbm_matches;

/* Message of the last failing call on this thread ("" if none). */
const char* sbm_last_error(void);
/* Library version string. */
const char* sbm_version(void);

/* ---- lifecycle -------------------------------------------------------------------------------
 * Detector(), Detector(T), Detector(num_features, T, weak, strong)       line2Dup.h:149-152
 * T == NULL selects the default pyramid {4, 8}; num_features <= 0 selects 63.             */
int sbm_detector_create(int num_features, const int* T, int num_levels, float weak_threshold,
                        float strong_threshold, sbm_detector** out);
/* The 16-orientation variant, test/ori_16bit_experiment/line2Dup_16bit_ori.cpp: Detector(),
 * Detector(T), Detector(num_features, T) (line2Dup_16bit_ori.cpp:1036-1056).  T == NULL selects {5, 8};
 * num_features <= 0 selects 63; weak and strong thresholds are the variant's fixed 10 and 55
 * (:500-512, 1051-1056).  Such a detector quantizes into 16 orientations over 180 degrees with 16-bit
 * spread values and the variant's compiled response table (:216-296, 542-700), takes 3-channel images only
 * (its gradient loop reads three channels, :319-367), refuses a template with fewer candidates than
 * num_features and selects exactly num_features (:167-210, 479-481), allows up to 16383 features per
 * level (:1166), and has no addTemplate_rotate.  Every other entry point works on it unchanged. */
int sbm_detector_create16(int num_features, const int* T, int num_levels, sbm_detector** out);
/* 8, or 16 for a detector made by sbm_detector_create16. */
int sbm_orientation_bins(const sbm_detector* det, int* out);
int sbm_detector_destroy(sbm_detector* det);

/* ---- getters (line2Dup.h:163-175) ---- */
int sbm_pyramid_levels(const sbm_detector* det, int* out);                       /* pyramidLevels() */
int sbm_get_T(const sbm_detector* det, int level, int* out);                     /* getT() */
int sbm_get_modality(const sbm_detector* det, float* weak, int* num_features,
                     float* strong);                                              /* getModalities() */
int sbm_num_templates(const sbm_detector* det, const char* class_id, int* out);  /* numTemplates([id]) */
int sbm_num_classes(const sbm_detector* det, int* out);                          /* numClasses() */
/* classIds(): copies the index-th class id (sorted order) into buf (NUL-terminated). */
int sbm_class_id(const sbm_detector* det, int index, char* buf, int cap);
/* getTemplates(class_id, template_id)[level]: geometry, then features (x, y, label) and theta. */
int sbm_get_template(const sbm_detector* det, const char* class_id, int template_id, int level, int* width,
                     int* height, int* tl_x, int* tl_y, int* pyramid_level, int* num_features);
int sbm_get_template_features(const sbm_detector* det, const char* class_id, int template_id, int level,
                              int32_t* xyl, float* theta, int cap);

/* ---- bank building ----------------------------------------------------------------------------
 * addTemplate(sources, class_id, object_mask, num_features)             line2Dup.h:158-159
 * template_id receives the new id, or -1 when a level has too few features (the reference's
 * "too few features, abort", line2Dup.cpp:487-490); the call itself then still returns SBM_OK. */
int sbm_add_template(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
                     const uint8_t* mask, const char* class_id, int num_features, int* template_id);
/* addTemplate_rotate(class_id, zero_id, theta, center)                  line2Dup.h:161 */
int sbm_add_template_rotate(sbm_detector* det, const char* class_id, int zero_id, float theta, float center_x,
                            float center_y, int* template_id);
/* A template pyramid given directly (what readClass builds from YAML, line2Dup.cpp:1436-1452):
 * tinfo6 holds per level width, height, tl_x, tl_y, num_features, feature_offset; xyl holds
 * (x, y, label) triples; theta may be NULL. */
int sbm_add_template_features(sbm_detector* det, const char* class_id, int num_levels, const int32_t* tinfo6,
                              const int32_t* xyl, const float* theta, int* template_id);

/* ---- matching ---------------------------------------------------------------------------------
 * match(sources, threshold, class_ids, masks)                            line2Dup.h:154-156
 * class_ids == NULL or num_class_ids == 0 matches every class.  Results are in the canonical
 * order: similarity desc, template_id asc, class asc, y asc, x asc, then std::unique with
 * Match::operator== (line2Dup.cpp:1113-1116, line2Dup.h:117-129). */
int sbm_match(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
              float threshold, const char* const* class_ids, int num_class_ids, const uint8_t* mask, int mode,
              sbm_matches** out);
/* Same as sbm_match with the image (and mask) already in device memory: d_img / d_mask are device
 * pointers on the detector's GPU; the mask is rows*cols contiguous bytes. */
int sbm_match_device(sbm_detector* det, const uint8_t* d_img, int rows, int cols, int stride, int channels,
                     float threshold, const char* const* class_ids, int num_class_ids, const uint8_t* d_mask,
                     int mode, sbm_matches** out);
/* sbm_match_device that allocates nothing: no reference counterpart (the reference allocates its
 * matrices per call).  For callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Every per-image buffer (level images, quantized
 * orientations, linear memories, candidate lists, counters) is carved from one fixed 64 MiB device array
 * inside the library, one per GPU, held until the process ends; the coarse stage holds up to 65,536
 * candidates.  The bank must be in HBM already (sbm_upload_bank after its last change), else SBM_EINVAL;
 * an image or selection that needs more than the array (about 2048 x 2048 for a gray image), or a coarse
 * stage passing more candidates, fails with SBM_EBUDGET.  The matches equal sbm_match_device's in batched
 * mode: xyct receives (x, y, class_index, template_id) and similarity one float for the first min(cap,
 * count) of them, in the canonical order; *count = all matches; class_index indexes sbm_class_id's list.
 * Calls on one GPU share its array (each call clears what it needs): calls that overlap on one GPU race,
 * and the caller serializes them. */
int sbm_match_device_noalloc(sbm_detector* det, const uint8_t* d_img, int rows, int cols, int stride, int channels,
                             float threshold, const char* const* class_ids, int num_class_ids, const uint8_t* d_mask,
                             int32_t* xyct, float* similarity, int cap, int* count);
/* An image (and optional mask) copied once into HBM, for repeated matching without host transfers. */
typedef struct sbm_image sbm_image;
int sbm_image_create(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
                     const uint8_t* mask, sbm_image** out);
int sbm_image_destroy(sbm_image* im);
int sbm_match_image(sbm_detector* det, const sbm_image* im, float threshold, const char* const* class_ids,
                    int num_class_ids, int mode, sbm_matches** out);
/* Several HBM-resident images in batched mode: images of one size bucket go out in launches of several
 * images (sbm_set_images_per_launch), each through the captured graph of the whole pipeline for its size and
 * image count (linear memories stay resident), and `slots` streams keep that many launches in flight.
 * outs[i] receives image i's matches, exactly as sbm_match_image would return them. */
int sbm_match_batch(sbm_detector* det, const sbm_image* const* images, int n, float threshold,
                    const char* const* class_ids, int num_class_ids, int slots, sbm_matches** outs);
/* sbm_match_batch with the match lists written into the caller's arrays instead of one sbm_matches per image:
 * image i's counts[i] matches follow image i-1's, as (x, y, class_index, template_id) in xyct and one float in
 * similarity, until cap matches in all have been written (counts[] holds every list's full length; a sum
 * above cap means a call with a larger cap returns the rest).  class_index refers to sbm_class_id's list. */
int sbm_match_batch_arrays(sbm_detector* det, const sbm_image* const* images, int n, float threshold,
                           const char* const* class_ids, int num_class_ids, int slots, int32_t* xyct,
                           float* similarity, int cap, int* counts);
/* sbm_match_batch_arrays for images in host memory (the reference's callers hand over host images,
 * line2Dup.h:154-156): image i is imgs[i] (rows[i] x cols[i], row stride strides[i] bytes, channels[i]
 * channels, mask masks[i] or none; masks may be NULL).  The images of a launch are staged through pinned
 * memory and copied with one DMA per launch; launches on different slots overlap their copies with each
 * other's kernels. */
int sbm_match_batch_host_arrays(sbm_detector* det, const uint8_t* const* imgs, const int* rows, const int* cols,
                                const int* strides, const int* channels, const uint8_t* const* masks, int n,
                                float threshold, const char* const* class_ids, int num_class_ids, int slots,
                                int32_t* xyct, float* similarity, int cap, int* counts);
/* Images per launch of sbm_match_batch (no reference counterpart): 1..16 images of one size bucket share every
 * launch of the pipeline -- the preprocessing kernels cover all of them, one coarse launch takes (image, tile,
 * template) items over the whole bank, one refinement launch per level takes every image's candidates; 0
 * (the default) puts half of each bucket in a launch, so two launches alternate on the slots. */
int sbm_set_images_per_launch(sbm_detector* det, int images_per_launch);
/* Waits for all work queued on the detector's stream. */
int sbm_synchronize(sbm_detector* det);
int sbm_matches_count(const sbm_matches* m, int* count);
/* Match{x, y, similarity, class_id, template_id} (line2Dup.h:108-141); class as an index into
 * sbm_matches_class_id. */
int sbm_matches_get(const sbm_matches* m, int i, int* x, int* y, float* similarity, int* class_index,
                    int* template_id);
/* Bulk copy: xyct receives (x, y, class_index, template_id) per match. */
int sbm_matches_copy(const sbm_matches* m, int32_t* xyct, float* similarity, int cap, int* copied);
int sbm_matches_class_id(const sbm_matches* m, int class_index, char* buf, int cap);
int sbm_matches_free(sbm_matches* m);

/* ---- shapeInfo_producer (line2Dup.h:209-332) ----
 * transform(src, angle, scale) (line2Dup.h:245-253): cv::getRotationMatrix2D((cols/2, rows/2), angle,
 * scale) then cv::warpAffine, INTER_LINEAR, zero border, same size; bit-exact with OpenCV's fixed-point
 * warp.  binary_mask != 0 gives mask_of's "transform(mask) > 0" as 0/255 (line2Dup.h:326-328).
 * det may be NULL: the warp then runs on the current device. */
int sbm_shape_transform(sbm_detector* det, const uint8_t* src, int rows, int cols, int stride, int channels,
                        float angle, float scale, int binary_mask, uint8_t* dst);
/* produce_infos (line2Dup.h:281-320): writes min(cap, count) (angle, scale) pairs; *count = all. */
int sbm_produce_infos(const float* angle_range, int n_angle_range, const float* scale_range, int n_scale_range,
                      float angle_step, float scale_step, float eps, float* angles, float* scales, int cap, int* count);
/* save_infos / load_infos (line2Dup.h:255-279): FileStorage YAML "infos: [{angle, scale}...]". */
int sbm_save_infos(const char* path, const float* angles, const float* scales, int n);
int sbm_load_infos(const char* path, float* angles, float* scales, int cap, int* count);
/* addTemplate(shapes.src_of(info), class_id, shapes.mask_of(info), num_features[i]) for n infos in order
 * (line2Dup.cpp:1270-1304 per info; test.cpp:242-292, 390-405 are the reference's loops).  mask may be
 * NULL (all 255, as shapeInfo_producer makes it); num_features may be NULL or hold 0s (the detector's
 * default).  template_ids[i] = the new id, or -1 where the reference's addTemplate returns -1.
 * Warps and gradients run on the GPU; feature selection on host threads. */
int sbm_add_templates_transformed(sbm_detector* det, const uint8_t* src, int rows, int cols, int stride, int channels,
                                  const uint8_t* mask, const char* class_id, const float* angles, const float* scales,
                                  const int* num_features, int n, int* template_ids);

/* ---- serving (no reference counterpart): a continuous-batching request loop over the detector's
 * resident bank.  `slots` pipelines (streams) run admitted requests concurrently.  `hbm_budget_bytes`
 * covers the detector's resident bank and every slot's buffers (the detector's own match buffers are given
 * back when the server is created; they grow again if the detector matches after the server is gone).  A
 * request is admitted when a slot is free and the bank plus the slots, with that slot's memory for the
 * request's image size, fit the budget (idle slots give their memory back when that makes room), and is
 * deferred otherwise.  Queued requests that match the admitted one in size, channels, mask presence,
 * threshold and selection join its launch (up to 8, within the budget): every kernel of the launch covers
 * them all.  Requests larger than the whole budget fail with SBM_EBUDGET; malformed requests are refused at
 * submission.  Each result equals what sbm_match (batched) returns for that image alone.  The detector's
 * bank must not be modified, nor the detector used, while a server is attached to it. */
#define SBM_ETIMEOUT (-11)
typedef struct sbm_server sbm_server;
int sbm_server_create(sbm_detector* det, int slots, long long hbm_budget_bytes, sbm_server** out);
int sbm_server_destroy(sbm_server* srv);
int sbm_server_submit(sbm_server* srv, const uint8_t* img, int rows, int cols, int stride, int channels,
                      const uint8_t* mask, float threshold, const char* const* class_ids, int num_class_ids,
                      long long* request_id);
/* Waits (timeout_ms < 0: forever) for a request; returns SBM_OK with its matches, the request's own
 * error code, or SBM_ETIMEOUT.  latency_ms: submission to completion. */
int sbm_server_wait(sbm_server* srv, long long request_id, int timeout_ms, sbm_matches** out, double* latency_ms);
/* peak_bytes: the most device memory the bank and the slots held together (at most the budget). */
int sbm_server_stats(sbm_server* srv, long long* submitted, long long* admitted, long long* completed,
                     long long* failed, long long* deferrals, long long* peak_bytes);
/* Slot launches so far: queued requests of one size, channels, mask presence, threshold and selection share
 * a launch (up to 8, as far as the budget allows), so admitted / launches is the mean launch size. */
int sbm_server_launches(sbm_server* srv, long long* launches);

/* ---- persistence (line2Dup.h:177-185) ---- */
int sbm_read_classes(sbm_detector* det, const char* const* class_ids, int n, const char* format); /* readClasses */
int sbm_write_classes(const sbm_detector* det, const char* format);                               /* writeClasses */
/* readClass(fn, class_id_override) on a YAML file; the class id read is copied to out. */
int sbm_read_class_file(sbm_detector* det, const char* path, const char* class_id_override, char* out, int cap);
int sbm_write_class_file(const sbm_detector* det, const char* class_id, const char* path);       /* writeClass */
int sbm_write_detector_file(const sbm_detector* det, const char* path);                           /* write() */
int sbm_read_detector_file(sbm_detector* det, const char* path);                                  /* read() */
/* Packed binary bank file: the whole bank (every class, its templates, features and feature angles) and
 * the detector's thresholds and num_features in one file, for a serving detector's start-up without YAML
 * parsing.  No reference counterpart: it sits beside readClasses / writeClasses (line2Dup.h:183-185),
 * whose YAML files stay the interchange format.  ".gz" paths are gzip-compressed.  Loading needs the
 * same orientation count and pyramid (levels and T) as the saving detector (else SBM_EINVAL), adds the
 * file's classes (SBM_EDUPCLASS if one is present; nothing is added then), restores the thresholds and
 * num_features, and reports the classes added in *num_classes (may be NULL).  SBM_EIO on a bad file. */
int sbm_save_bank(const sbm_detector* det, const char* path);
int sbm_load_bank(sbm_detector* det, const char* path, int* num_classes);
/* Packs the bank and copies it to HBM now, instead of at the next match (no reference counterpart: the
 * reference's bank lives in host memory).  A serving start-up pays the upload here; the bank is packed
 * again only after it changes. */
int sbm_upload_bank(sbm_detector* det);

/* ---- post-match NMS: cv_dnn::NMSBoxes(boxes, scores, score_thr, nms_thr, indices, eta, top_k)
 *                                                                        test.cpp:114-119
 * boxes: (x, y, w, h) int32 quadruples; indices receives the kept box indices. */
int sbm_nms_boxes(const int32_t* boxes, const float* scores, int n, float score_threshold, float nms_threshold,
                  float eta, int top_k, int32_t* indices, int* count);
/* match(...) followed by the demos' NMS step on the same matches (test.cpp:428-451):
 * cv_dnn::NMSBoxes over boxes (x, y, templ[0].width, templ[0].height) with their similarities, eta 1.
 * The batched matcher runs, then one device kernel builds the canonical match list (sort, unique) and
 * the kept set, so only the results cross PCIe; lists longer than 2048 entries are finished on the host
 * with the same result.  *kept receives the kept matches in NMS order (the reference's `indices` mapped
 * through its match list); *all (may be NULL) receives the full list, as sbm_match would return it. */
int sbm_match_nms(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
                  const uint8_t* mask, float threshold, const char* const* class_ids, int num_class_ids,
                  float score_threshold, float nms_threshold, int top_k, sbm_matches** kept, sbm_matches** all);
/* Same with an HBM-resident image. */
int sbm_match_image_nms(sbm_detector* det, const sbm_image* im, float threshold, const char* const* class_ids,
                        int num_class_ids, float score_threshold, float nms_threshold, int top_k,
                        sbm_matches** kept, sbm_matches** all);

/* ---- tuning and instrumentation (no reference counterpart) ---- */
/* Coarse-kernel tile shape: positions per lane (4, 8, 16; 32 with gather 3), wavefronts per block,
 * persistent blocks per CU, the gather (0 = aligned dword loads realigned in registers, 1 = unaligned
 * buffer vector loads, 2 = aligned vector loads + the next lane's data by DPP, 3 = as 2 on the
 * nibble-packed coarse level) and split (1, 2, 4, 8: wavefronts of one block sharing a work item's
 * features; waves_per_block is ignored when split > 1). */
int sbm_set_tuning(sbm_detector* det, int positions_per_lane, int waves_per_block, int blocks_per_cu, int gather,
                   int split);
int sbm_get_tuning(const sbm_detector* det, int* positions_per_lane, int* waves_per_block, int* blocks_per_cu,
                   int* gather, int* split);
/* Autotuner: on this image, threshold and class selection it times the coarse stage on the path the device
 * takes -- the gather kernel (plus the nibble packing it needs) over its 240 tile shapes, or the
 * nonzero-response scatter over its 17 (threads, score-map segment, blocks per CU) and then the split of its pruned
 * scatter (100, 60, 50, 40 percent) -- then the refinement over
 * its 140 (wavefronts per candidate and blocks per CU in 14 pairs, loads in flight, 5 layouts, each with the
 * copy it reads); checks that every shape returns the same
 * matches, and keeps the fastest of each stage in the detector.  With sbm_set_images_per_launch n > 1 it
 * times launches of n copies of the image as a stream (two launches in flight, host time per launch) on the path
 * such launches take; when they take the list it also times the gather kernel's shapes on them and, if that is
 * faster, sets coarse mode 0 (sbm_get_coarse_mode), else searches the gather shapes on the image alone when only
 * the launches take the list.  Stream timings drift by about as much as the best shapes differ, so the four
 * fastest shapes of each stage timed as a stream are timed twice more in turn, and where the best list shape and
 * the best gather shape come within a tenth of each other both are timed three more times before the path is
 * chosen; a shape's fastest timing counts.  The outputs give the gather kernel's shape;
 * trace (may be NULL) receives up to trace_cap of its rows (ppl, wpb, bpc, gather, split, microseconds);
 * sbm_autotune_row reads every stage's rows. */
int sbm_autotune(sbm_detector* det, const sbm_image* im, float threshold, const char* const* class_ids,
                 int num_class_ids, int iters, int* positions_per_lane, int* waves_per_block, int* blocks_per_cu,
                 int* gather, int* split, int32_t* trace, int trace_cap, int* trace_count);
/* The autotuner on several resident images of one size and channel count, none masked: single-image stages run
 * on images[0], and with sbm_set_images_per_launch k > 1 the k-image launches take images[0..n) in turn instead
 * of copies of one image, so that the choices (the coarse path among them) follow a stream of distinct scenes,
 * each launch's lists checked against its scenes matched alone.  No counterpart in the reference (line2Dup.cpp has no
 * launch shapes); the shapes are read back with sbm_get_tuning, sbm_get_tuning_param and sbm_autotune_row. */
int sbm_autotune_images(sbm_detector* det, const sbm_image* const* images, int n, float threshold,
                        const char* const* class_ids, int num_class_ids, int iters);
/* Row `index` of the last autotune: stage 0 (gather kernel: ppl, wpb, bpc, gather, split), 1 (nonzero-response
 * scatter: threads, segment, blocks per CU, order, split) or 2 (refinement: waves, group, blocks per CU, layout, 0), five
 * params and the stage's time in microseconds.  *count (may be NULL) = the rows; index < 0 only counts. */
int sbm_autotune_row(const sbm_detector* det, int index, int* stage, int32_t* params, float* us, int* count);
/* One named launch-shape parameter: "positions_per_lane", "waves_per_block", "blocks_per_cu", "gather", "split"
 * (the gather coarse kernel, as sbm_set_tuning), "list_threads" (256, 512, 1024), "list_segment" (4096, 8192,
 * 16384), "list_blocks_per_cu" (1, 2, 4, 8; within a CU's 2048 threads and 160 KB of LDS, a score map taking 8 bytes per position), "list_order" (0: the
 * blocks take work items from a counter; 1: fixed shares, each XCD a contiguous eighth), "list_split" (10..100:
 * the percent of each template's features scattered before the positions that can still pass are kept and the
 * remaining features' responses gathered there; 100 scatters them all) for the nonzero-response scatter, "refine_waves" (1, 2, 4, 8, 16), "refine_group" (8, 16), "refine_blocks_per_cu" (1..32),
 * "refine_layout" (0: the refinement reads the linear memories, 8 bytes per lane; 1: a banded copy of each
 * refined level, made after its linear memories, in which every 16 x 16 window is one run of 512 bytes; 2: the
 * linear memories, 16 bytes per lane and two features per load; 3: the banded copy, one window column per
 * 16-byte lane load and four features per load; 4: a nibble-banded copy, two rows per byte, in which every
 * window column is one 16-byte slot and every window one run of 256 bytes, four features per load) for the
 * refinement.  Every shape returns the same matches. */
int sbm_set_tuning_param(sbm_detector* det, const char* name, int value);
int sbm_get_tuning_param(const sbm_detector* det, const char* name, int* value);
/* Per-stage device times of the last match (HIP events on the detector's stream) when enabled. */
/* Batched coarse stage: 0 = the gather kernel above; 1 (default) = the nonzero-response path when at
 * most 5% of the coarse level's responses are nonzero, else the gather kernel; 2 = the nonzero-response
 * path whenever its list fits (one response in eight).  Both give the same matches; the choice is made
 * on the device for every image. */
int sbm_set_coarse_mode(sbm_detector* det, int mode);
int sbm_get_coarse_mode(const sbm_detector* det, int* mode);
/* Path of the last batched match's coarse stage (1 = nonzero-response list) and the number of nonzero
 * responses on its coarse level (-1 when not listed). */
int sbm_last_coarse_path(const sbm_detector* det, int* sparse, int* nonzero);
/* Features of the candidates the last match's finest-level refinement scored, summed (the refinement's
 * algorithmic work: 256 positions per feature per candidate). */
int sbm_last_refine_features(const sbm_detector* det, long long* features);
/* List entries the last batched match's nonzero-response path read (0 when it did not run). */
int sbm_last_coarse_entries(const sbm_detector* det, long long* entries);
/* enabled: 0 off; 1 per-kernel HIP events (matches run eagerly, kernel by kernel; so do the launches of
 * sbm_match_batch*, each of which then times its own kernels); 2 only the device span of each replayed graph
 * (input copy to result copy), read with sbm_last_device_ms. */
int sbm_set_timing(sbm_detector* det, int enabled);
/* Device time of the last match: the replayed graph's span (timing 2), or the sum of the timed
 * stages (timing 1). */
int sbm_last_device_ms(const sbm_detector* det, float* ms);
int sbm_last_timing(const sbm_detector* det, float* preprocess_ms, float* coarse_ms, float* refine_ms,
                    int* coarse_candidates, int* final_candidates);
/* Per-kernel device times of the last match (timing enabled), or of the last launch of the last
 * sbm_match_batch* call (every kernel of that launch covering all of its images); the sbm_last_* counters
 * describe the same match or launch.  *count receives the number of timed launches; for index < *count,
 * name (e.g. "gradient_l0", "coarse_batched") and ms are filled. */
int sbm_last_kernel_time(const sbm_detector* det, int index, char* name, int cap, float* ms, int* count);
/* Algorithmic gather bytes of the coarse stage for an image size and class selection:
 * sum over templates of (in-bounds lowest-level features) x (template positions). */
int sbm_coarse_gather_bytes(sbm_detector* det, int rows, int cols, const char* const* class_ids, int num_class_ids,
                            long long* bytes);
/* Runs one plain device-to-device copy kernel of `bytes` bytes on the detector's stream (a known-size
 * reference for scaling hardware traffic counters); the buffers are freed before returning. */
int sbm_copy_probe(sbm_detector* det, long long bytes);
/* Device memory held by the detector, in bytes. */
int sbm_device_bytes(const sbm_detector* det, long long* bytes);

/* ---- stage access for parity tests (no reference counterpart; they expose the intermediates the
 * reference keeps inside match: quantized orientations, linear memories, similarity maps) ---- */
int sbm_prepare_image(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
                      const uint8_t* mask);
/* what: 0 = masked quantized orientations (rows_l*cols_l bytes; uint16 each with 16 orientations),
 * 1 = linear memories [planes][T*T][H*W] (planes*rows_l*cols_l bytes, planes = 8 or 16),
 * 2 = the level's image (rows_l*cols_l*channels bytes), 3 = the coarse level's nibble-packed linear memories
 * (with gather 3), 4 = a refined level's banded copy (with refine_layout 1: planes*T*T*ceil(H/16)*(W+16)*32
 * bytes, then 1024 zero bytes) */
int sbm_read_level(sbm_detector* det, int level, int what, uint8_t* out);
/* similarity() map of a template at the lowest level of the prepared image (H*W uint16). */
int sbm_similarity_map(sbm_detector* det, const char* class_id, int template_id, uint16_t* out);
/* similarityLocal() 16x16 map of a template at a level around center (cx, cy). */
int sbm_similarity_local(sbm_detector* det, const char* class_id, int template_id, int level, int cx, int cy,
                         uint16_t* out);
/* quantizedOrientations() of one image: quantized (unmasked; rows*cols bytes, or rows*cols uint16 with
 * 16 orientations), magnitude and angle (may be NULL). */
int sbm_gradient(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels, uint8_t* q,
                 float* magnitude, float* angle);
/* The device's (dx, dy) -> 8-bin orientation table for dx, dy in [lo, hi], within [-1020, 1020] (the Sobel
 * responses of 8-bit images). */
int sbm_bin_table(sbm_detector* det, int lo, int hi, uint8_t* out);
/* Self-check of the gradient kernels' quotient step (no counterpart in the reference): *mismatches = the integer
 * pairs 0 <= num <= den <= 1020 whose quotient num / (den + DBL_EPSILON) differs from the IEEE single-precision
 * one on this GPU; 0 is required for the phase to equal OpenCV's fastAtan2. */
int sbm_division_check(sbm_detector* det, int* mismatches);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif
#endif /* SBM_HIP_H */
