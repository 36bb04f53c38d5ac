/*
 * vj_api.h -- C ABI of the MI355X-native Viola-Jones hot path (libvj_hip.so).
 *
 * Plain pointers and sizes only; no C++ or HIP types cross this boundary
 * (streams and events are opaque void* that are hipStream_t / hipEvent_t).
 * Each entry point names the reference interface it replaces, as
 * path:line in the reference repository (hengyizhu/CS677s26).
 *
 * Ownership: input pointers are borrowed for the duration of the call;
 * handles own their device and pinned-host buffers.  No C++ exception
 * crosses the ABI; failures return a status code and vj_last_error() gives
 * a message (thread-local).
 *
 * Device pointers ("d" prefix) must come from vj_malloc (or any hipMalloc'd
 * memory of the current device).  Host pointers ("h" prefix) are ordinary
 * memory.
 */
#ifndef VJ_API_H
#define VJ_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* vj::Status, include/vj/common_types.cuh:11-17 (same numeric values;
 * VJ_DEVICE_ERROR is the reference's kCudaError slot). */
enum {
    VJ_OK = 0,
    VJ_INVALID_ARG = 1,
    VJ_DEVICE_ERROR = 2,
    VJ_OUT_OF_MEMORY = 3,
    VJ_RUNTIME_ERROR = 4
};

/* POD records, byte-identical to include/vj/common_types.cuh:26-143. */
typedef struct vj_rect { int32_t p0, p1, p2, p3; float w; } vj_rect;               /* FastRect     */
typedef struct vj_feature {                                                           /* HaarFeature  */
    vj_rect r[3];
    uint8_t rectCount, tilted;
    uint16_t reserved;
} vj_feature;
typedef struct vj_stump { int32_t feature; float theta, left, right; } vj_stump;     /* GpuStump4 (x = int bits) */
typedef struct vj_stage { int32_t first, ntrees; float threshold; } vj_stage;        /* CascadeStage */
typedef struct vj_detection { int32_t x, y, w, h, scaleIdx; float score; } vj_detection; /* Detection */
typedef struct vj_weak {                                                              /* WeakClassifier + round stats */
    int32_t feature;        /* chosen feature index (-1: none)                  */
    float theta;            /* 0.5f*(v_k + v_k+1)                                 */
    int32_t parity;         /* +1: resp<theta votes -1; -1: resp<theta votes +1 */
    int32_t split;          /* position k in the feature's sorted order          */
    double err;             /* weighted error (exact sum of the search weights)  */
    double alpha;           /* C = log((1-e')/e'), e' = clamp(err, 1e-6, 1-1e-6) */
    float left, right;      /* leaf values stored in the stump (float)(-+C)      */
    float stageThreshold;   /* calibrated threshold after this round             */
    float hitRate, falseAlarm;
    int32_t posPass, negSurvive;
    int32_t stageSatisfied; /* hit >= minHitRate && fa <= maxFalseAlarm, or no negatives left */
} vj_weak;

typedef void* vj_stream; /* hipStream_t; NULL = the default (null) stream */
typedef void* vj_event;  /* hipEvent_t */

const char* vj_last_error(void);

/* ---- device plumbing: the reference calls the CUDA runtime directly
 * (src/main.cu:52-101 CudaBuffer; cudaMemcpy/cudaEvent* throughout). ------- */
int vj_device_count(int* n);
int vj_device_name(char* buf, int cap);
int vj_malloc(void** dptr, size_t bytes);
int vj_free(void* dptr);
int vj_host_alloc(void** hptr, size_t bytes); /* pinned */
int vj_host_free(void* hptr);
int vj_memcpy_htod(void* dst, const void* src, size_t bytes, vj_stream s);
int vj_memcpy_dtoh(void* dst, const void* src, size_t bytes, vj_stream s);
int vj_memcpy_dtod(void* dst, const void* src, size_t bytes, vj_stream s);
int vj_memset(void* dst, int value, size_t bytes, vj_stream s);
int vj_stream_create(vj_stream* s);
int vj_stream_destroy(vj_stream s);
int vj_stream_synchronize(vj_stream s);
int vj_device_synchronize(void);
int vj_event_create(vj_event* e);
int vj_event_destroy(vj_event e);
int vj_event_record(vj_event e, vj_stream s);
int vj_event_synchronize(vj_event e);
int vj_event_elapsed_ms(vj_event start, vj_event stop, float* ms);

/* ---- Haar pool: FaceVisionEngine::buildHaarFeaturePool / countHaarFeatures,
 * src/face_vision_engine.cu:877-1020.  mode 0 Basic, 1 Core, 2 All.
 * *count receives the pool size; at most cap records are written. -------- */
int vj_haar_pool(int winW, int winH, int mode, vj_feature* hOut, int cap, int* count);

/* ---- model file: ModelIO::saveCascadeModel / loadCascadeModel,
 * src/model_io.cpp:66-167 (.bin "VJCBIN1", version 1, byte-identical). --- */
int vj_model_save(const char* path, int winW, int winH,
                  const vj_feature* hFeatures, int64_t nFeatures,
                  const vj_stump* hStumps, int64_t nStumps,
                  const vj_stage* hStages, int64_t nStages);
int vj_model_peek(const char* path, int* winW, int* winH,
                  int64_t* nFeatures, int64_t* nStumps, int64_t* nStages);
int vj_model_load(const char* path, vj_feature* hFeatures, int64_t capFeatures,
                  vj_stump* hStumps, int64_t capStumps, vj_stage* hStages, int64_t capStages);

/* ---- grouping: FaceVisionEngine::groupRectanglesCpu,
 * src/face_vision_engine.cu:1401-1539 (host).  out needs room for n. ---- */
int vj_group_rectangles(const vj_detection* hIn, int n, int minNeighbors, float eps,
                        vj_detection* hOut, int* count);
/* The same for nImages boxes lists at once (hIn[i] with n[i] boxes into
 * hOut[i], which needs room for n[i]; count[i] set), each result equal to
 * vj_group_rectangles on that list; the work of all lists is spread over the
 * library's host threads.  No reference counterpart (batched detect uses it). */
int vj_group_rectangles_many(const vj_detection* const* hIn, const int* n, int nImages, int minNeighbors,
                             float eps, vj_detection* const* hOut, int* count);

/* ---- PGM / PPM: loadPGM / savePPM, src/main.cu:286-371 -------------------- */
int vj_pgm_peek(const char* path, int* w, int* h);
int vj_pgm_load(const char* path, uint8_t* hOut, size_t cap);
int vj_ppm_save_gray_boxes(const char* path, const uint8_t* hGray, int w, int h,
                           const vj_detection* hBoxes, int n);

/* ---- raw sample files: loadRawSampleBin, src/main.cu:122-162 ------------
 * A file of patches of samplePixels bytes each, back to back.  *count is the
 * number of patches.  An unreadable or empty file, or one whose size is not
 * a whole number of patches, gives VJ_INVALID_ARG and the reference's
 * message ("Cannot open: <path>", "Empty file: <path>", "File size is not
 * divisible by samplePixels: <path>").  vj_samples_load needs cap >= the
 * file size. */
int vj_samples_count(const char* path, int samplePixels, int* count);
int vj_samples_load(const char* path, int samplePixels, uint8_t* hOut, size_t cap, int* count);

/* ---- PGM files under a directory: collectPgmPaths, src/main.cu:226-244 --
 * Every regular file named *.pgm (extension in any case) under dir, at any
 * depth, as NUL-terminated paths packed into buf in ascending byte order
 * (the reference keeps directory order; sorted, a seed picks the same
 * images on every file system).  *bytes is the size the whole list needs,
 * *count the number of paths; buf is filled only when cap >= *bytes.  A
 * missing directory gives an empty list. */
int vj_pgm_list(const char* dir, char* buf, size_t cap, size_t* bytes, int* count);

/* ---- hard-negative crops: cropResizeTo24, src/main.cu:245-284 (host) -----
 * The box (x, y, w, h) of a gray image (row pitch stride) resampled to
 * outW x outH by pixel-centre bilinear interpolation with clamped taps (the
 * reference's window is 24 x 24). ---- */
int vj_crop_resize(const uint8_t* hGray, int w, int h, int stride, const vj_detection* box,
                   int outW, int outH, uint8_t* hOut);

/* ==== FaceVisionEngine, include/vj/face_vision_engine.cuh:19-196 ======== */
typedef struct vj_engine vj_engine;

typedef struct vj_integral_view {  /* IntegralImageSetT, common_types.cuh:48-60 */
    int32_t nSamples, iiWidth, iiHeight, iiArea;
    const uint32_t* sumT;          /* [iiArea][nSamples] (row-major image when nSamples == 1) */
    const uint64_t* sqsumT;        /* [iiArea][nSamples], full 64-bit squares                 */
    const float* invNorm;          /* [nSamples]                                               */
} vj_integral_view;

/* FaceVisionEngine ctor, src/face_vision_engine.cu:644-725 (throws there). */
int vj_engine_create(int winW, int winH, int maxSamples, int maxImageW, int maxImageH,
                     int enableTilted, int maxDetections, vj_engine** out);
void vj_engine_destroy(vj_engine* e);
/* uploadFeaturePool, src/face_vision_engine.cu:1022-1055 */
int vj_engine_upload_feature_pool(vj_engine* e, const vj_feature* hFeatures, int n, vj_stream s);
/* deviceFeaturePool / featurePoolSize, include/vj/face_vision_engine.cuh:62-63 */
int vj_engine_feature_pool(vj_engine* e, const vj_feature** dFeatures, int* n);
/* computeIntegralBatchTransposed, src/face_vision_engine.cu:1057-1148 */
int vj_engine_integral_batch(vj_engine* e, const uint8_t* dGray, int w, int h, int stride,
                             int nSamples, int computeInvNorm, vj_stream s);
/* integralSetView, src/face_vision_engine.cu:1150-1162 */
int vj_engine_integral_view(vj_engine* e, vj_integral_view* out);
/* One gray image's integral without allocating.  No counterpart in the
 * reference: it is the single-image form of vj_engine_integral_batch
 * (computeIntegralBatchTransposed with n = 1) for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation.  It calls no hipMalloc, hipFree
 * or other allocating routine: the caller passes the image and the outputs,
 * dSum (u32) and dSq (u64), (h+1) x (w+1) each, row-major with row pitch
 * w+1, row 0 and column 0 zero.  The intermediate sums live in a fixed
 * device array of the library (under 3 MiB per GPU, held until the process
 * ends), sized for sides up to 8192 with w * h * 255 < 2^31; larger images
 * are refused with VJ_INVALID_ARG.  Every call on a GPU shares that array,
 * so calls that overlap on one GPU (other streams or host threads) race.
 * Asynchronous on s. */
int vj_integral_image_noalloc(const uint8_t* dGray, int w, int h, int stride, uint32_t* dSum, uint64_t* dSq,
                              vj_stream s);
/* setCascadeModel, src/face_vision_engine.cu:1176-1322 (copies its inputs). */
int vj_engine_set_cascade(vj_engine* e, const vj_feature* dFeatures, int nFeatures,
                          const vj_stump* dStumps, int nStumps,
                          const vj_stage* hStages, int nStages, vj_stream s);
/* clearCascadeModel, src/face_vision_engine.cu:1324-1354 */
int vj_engine_clear_cascade(vj_engine* e, vj_stream s);
/* detectMultiScale, src/face_vision_engine.cu:1541-1747.  Synchronises s.
 * Raw detections are in (level, y, x) scan order and truncated to the first
 * maxDetections of that order (the reference CPU path's order and cap).
 * *count receives the number of detections; min(*count, cap) are copied. */
int vj_engine_detect(vj_engine* e, const uint8_t* dImage, int w, int h, int stride,
                     float scaleFactor, int minNeighbors, int minObjectSize, int maxDetections,
                     int applyGrouping, vj_detection* hOut, int cap, int* count, vj_stream s);

typedef struct vj_detect_stats {
    int32_t levels;          /* pyramid levels scanned                    */
    int64_t windows;         /* windows visited (after step skipping)     */
    int64_t rawDetections;   /* before truncation and grouping            */
    int64_t pyramidPixels;   /* sum of level W*H                          */
    int64_t integralBytes;   /* algorithmic bytes of the integral kernels */
    int64_t cascadeBytes;    /* algorithmic bytes of the cascade kernel   */
    int32_t imagesGroupedOnHost; /* images whose grouping fell back to the
                                    host (pair buffer overflow, device limits) */
} vj_detect_stats;
/* Statistics of the engine's last detect (no reference counterpart). */
int vj_engine_last_stats(vj_engine* e, vj_detect_stats* out);
/* Per-phase device time of the last detect when timing is on (HIP events on
 * the detect's stream): ms[0] pyramid, ms[1] integrals, ms[2] cascade,
 * ms[3] compaction.  No reference counterpart (its detect-bench times the
 * whole call, src/main.cu:2471-2537). */
int vj_engine_set_timing(vj_engine* e, int on);
int vj_engine_phase_ms(vj_engine* e, float* ms);
/* Counting mode (no reference counterpart; instrumentation of the
 * cascade's work per stage, for its LDS-gather roofline): while on, detect
 * calls run cascade kernels that also count the windows entering each stage
 * (slower; never captured in a graph).  vj_engine_stage_windows gives those
 * counts for the last counted call (entry 0: the windows that pass the
 * variance gate), summed over the call's images; *nStages is the cascade's
 * stage count, out receives up to cap entries. */
int vj_engine_count_stages(vj_engine* e, int on);
int vj_engine_stage_windows(vj_engine* e, int64_t* out, int cap, int* nStages);

/* Batched detect (no reference counterpart: the reference detects one image
 * per call).  nImages images, each through the same pyramid, integral,
 * cascade and compaction kernels as vj_engine_detect but all images' work in
 * the same launches; the launch sequence of a batch shape that repeats is
 * captured in a HIP graph and replayed (vj_engine_set_graphs), with tile shapes from the
 * autotuner when it is on (vj_engine_set_autotune).  Per image the detections are identical to
 * vj_engine_detect on that image.  hOut receives the images' detections
 * concatenated in image order; hCounts[i] is image i's count; the return is
 * VJ_INVALID_ARG when the total exceeds cap. */
int vj_engine_detect_batch(vj_engine* e, int nImages, const uint8_t* const* dImages,
                           const int* w, const int* h, const int* stride,
                           float scaleFactor, int minNeighbors, int minObjectSize,
                           int maxDetections, int applyGrouping,
                           vj_detection* hOut, int cap, int* hCounts, vj_stream s);

/* Launch-path switches of the detect pipeline (no reference counterpart).
 * Graphs (default on): the launch sequence of a detect is captured in a HIP
 * graph on the second consecutive call of a plan shape and replayed while the
 * shape and the engine's buffers are unchanged; a shape seen once is launched
 * directly.  Autotune (default off): the first call of each batch shape
 * (image sizes, scale factor, minimum object size, cascade) times the
 * cascade kernels over a few candidate tile shapes, each with and without
 * the tail pass (without it the deep pass runs every remaining stage), and
 * keeps the fastest for that shape; turning it off restores the defaults.
 * Neither changes any detection.  vj_engine_tile_rows reports the cascade
 * tile rows in use for window steps 1, 2 and 3; vj_engine_cascade_split
 * reports {first deep-pass stage, first tail-pass stage, stages} in use (the
 * tail pass is off when the second equals the third). */
int vj_engine_set_graphs(vj_engine* e, int on);
int vj_engine_set_autotune(vj_engine* e, int on);
int vj_engine_tile_rows(vj_engine* e, int* rows3);
int vj_engine_cascade_split(vj_engine* e, int* split3);

/* Device bytes one detect of a w x h image occupies in an engine's work
 * buffers (input, pyramid, integrals and their scratch, window mask and
 * scores, output at min(maxDetections, the engine's cap)).  The served form
 * charges this against its HBM budget.  No reference counterpart. */
int vj_engine_detect_footprint(vj_engine* e, int w, int h, float scaleFactor, int minObjectSize,
                               int maxDetections, int64_t* bytes);

/* ==== Served detect (no reference counterpart) ===========================
 * A request loop with continuous batching: requests (host images of mixed
 * sizes) queue on submit; each of `slots` device slots (one engine and one
 * stream each, the cascade resident in every slot) takes as many queued
 * requests as fit its share of the HBM budget (hbmBudget / slots, charged by
 * vj_engine_detect_footprint) whenever it is free, and runs them as one
 * batched detect, so requests that arrive while others are on the GPU are
 * admitted as slots free up and the slots' work overlaps on the device.
 * Every request's detections are those of vj_engine_detect on its image
 * with the server's parameters.  A request larger than one slot's share is
 * refused at submit with VJ_OUT_OF_MEMORY.  Thread-safe: any thread may
 * submit and wait. */
typedef struct vj_server vj_server;
typedef struct vj_server_config {
    int32_t winW, winH;             /* cascade window                            */
    int32_t maxImageW, maxImageH;   /* largest request                           */
    int32_t slots;                  /* batches in flight on the device, >= 1     */
    int32_t maxBatch;               /* requests per batch, >= 1                  */
    int64_t hbmBudget;              /* device bytes of admitted request state    */
    float scaleFactor;              /* detectMultiScale parameters, per server   */
    int32_t minNeighbors, minObjectSize, maxDetections, applyGrouping;
} vj_server_config;
typedef struct vj_server_stats {
    int64_t submitted, completed, rejected, batches;
    int64_t peakBytesInFlight;      /* largest sum of admitted footprints        */
    int32_t maxSlotsBusy, maxBatchSeen;
} vj_server_stats;

/* The cascade is given as host arrays and uploaded once per slot. */
int vj_server_create(const vj_server_config* cfg, const vj_feature* hFeatures, int nFeatures,
                     const vj_stump* hStumps, int nStumps, const vj_stage* hStages, int nStages,
                     vj_server** out);
/* Copies the image (row stride `stride` bytes) and queues it; *ticket names
 * the request for vj_server_wait. */
int vj_server_submit(vj_server* srv, const uint8_t* hImage, int w, int h, int stride, int64_t* ticket);
/* Blocks until the request is done; copies min(*count, cap) detections and
 * gives the arrival-to-result latency in ms.  A ticket is waited once. */
int vj_server_wait(vj_server* srv, int64_t ticket, vj_detection* hOut, int cap, int* count,
                   double* latencyMs);
int vj_server_stats_get(vj_server* srv, vj_server_stats* out);
/* Finishes every queued request, then frees the slots. */
void vj_server_destroy(vj_server* srv);

/* ==== AdaBoostTrainer, include/vj/boosting_core.cuh:20-161 =============== */
typedef struct vj_trainer vj_trainer;

/* AdaBoostTrainer ctor, src/boosting_core.cu:466-548 (throws there). */
int vj_trainer_create(int winW, int winH, int maxSamples, int maxFeaturesPerTile, vj_trainer** out);
void vj_trainer_destroy(vj_trainer* t);
/* bindTrainingSet, src/boosting_core.cu:684-745 (non-owning views; n <= 65535). */
int vj_trainer_bind(vj_trainer* t, const vj_integral_view* ii, const int8_t* dLabel,
                    const uint8_t* dActive, const float* dWeight, int nSamples);
/* evaluateFeatureResponses, src/boosting_core.cu:747-787: dOut [count][n] f32. */
int vj_trainer_evaluate_responses(vj_trainer* t, const vj_feature* dFeatures, int begin, int count,
                                  float* dOut, vj_stream s);
/* sortSamplesPerFeature, src/boosting_core.cu:829-892: per-feature ascending
 * (value bits, index) order of dResp rows, written as plain u16 indices
 * (any bound sample count up to 65,535). */
int vj_trainer_sort_per_feature(vj_trainer* t, const float* dResp, int count, uint16_t* dOutSorted,
                                vj_stream s);
/* evaluateAndFindThreshold, src/boosting_core.cu:789-827 (FeatureBestSplit
 * records, 16 B each: err f32, theta f32, split i32, parity i8, pad); up to
 * 65,535 samples. */
int vj_trainer_find_threshold(vj_trainer* t, const float* dResp, const uint16_t* dSorted, int count,
                              void* dOutBest, vj_stream s);
/* selectBestInTile, src/boosting_core.cu:894-925 (TileBestCandidate, 16 B). */
int vj_trainer_select_best(vj_trainer* t, const void* dBest, int begin, int count,
                           const uint8_t* dUsedMask, void* dOut, vj_stream s);

/* ---- fused stage / round form of the training loop (src/main.cu:1936-2348).
 * The samples of a stage are fixed, so each feature's sample order is sorted
 * once per stage (stage_begin) and every boosting round is one exact
 * weighted prefix scan over the stored orders (round). ------------------- */
typedef struct vj_round_config {
    float minHitRate;      /* default 0.99  (src/main.cu:1386) */
    float maxFalseAlarm;   /* default 0.40  (src/main.cu:1387) */
} vj_round_config;

/* Stage start (src/main.cu:1936-1997): integrals and invNorm of the n
 * 24x24 patches (positives first, n <= 65,535), presort of every feature of dFeatures,
 * weights 0.5/numPos and 0.5/activeNeg normalised, strong scores 0, used
 * mask cleared.  hActive may be NULL (all active). */
int vj_trainer_stage_begin(vj_trainer* t, const uint8_t* dPatches, const int8_t* hLabels,
                           const uint8_t* hActive, int nSamples, int numPos,
                           const vj_feature* dFeatures, int nFeatures, vj_stream s);
/* One boosting round (src/main.cu:2005-2300): best stump over unused
 * features, weight update, normalisation and stage-threshold calibration. */
int vj_trainer_round(vj_trainer* t, const vj_round_config* cfg, vj_weak* out, vj_stream s);
/* Stage end (src/main.cu:2302-2348): negatives with strong >= threshold stay
 * active; returns their count in *remainNeg. */
int vj_trainer_stage_end(vj_trainer* t, float stageThreshold, int* remainNeg, vj_stream s);
/* Copies of the trainer's per-sample state (for checks and debugging). */
int vj_trainer_read_state(vj_trainer* t, double* hWeights, float* hStrong, uint8_t* hActive,
                          float* hInvNorm);
/* Presorted orders of features [begin, begin+count), n entries each, for
 * checking the presort.  In the LDS layout (up to 20,480 samples) an entry is
 * the sample index with bit 15 set when a valid split follows the position;
 * in the wide layout (always above 20,480 samples, or as the stage's
 * schedule picks) the entries are plain 16-bit indices and the flags are kept
 * apart. */
int vj_trainer_read_orders(vj_trainer* t, int begin, int count, uint16_t* hOut);
/* The valid-split flags of the same orders, one byte (0 or 1) per position:
 * fl(v_k + 1e-7f) < v_{k+1} (src/boosting_core.cu:395-405). */
int vj_trainer_read_splits(vj_trainer* t, int begin, int count, uint8_t* hOut);

/* Sort and scan schedule of a stage (no counterpart in the reference, whose
 * block sizes are fixed, src/boosting_core.cu:25-29).  Four ints:
 *   [0] wide   1: global-memory radix sort and split-mask rows (forced above
 *              20,480 samples), 0: the LDS radix sort with flags in the entries;
 *   [1] ahead  search: steps of 16-byte order loads in flight per lane (2, 4, 8);
 *   [2] blocks search: workgroups per CU (1, 2, 4);
 *   [3] packed search: packed 32-bit weights in LDS even where int64 ones fit.
 * Every schedule gives the same rounds bit for bit.  With autotune on, the
 * first stage_begin of each (samples / 1024, log2 features, window) bucket
 * picks the layout from a sample of 16,384 features (presort and one scan in
 * each layout, scaled to one presort plus 50 rounds of the stage), presorts
 * every feature in it, then times the stage's first two rounds under every scan
 * candidate from the same saved state (restored afterwards) and keeps the
 * fastest (the default unless another is faster by more than 3%) for the
 * process; later stages of the bucket reuse the pick.  set_schedule forces one (NULL: back to the default {0, 4, 2, 0})
 * and turns autotune off; vj_trainer_schedule reports the schedule of the
 * current stage. */
int vj_trainer_set_autotune(vj_trainer* t, int on);
/* No counterpart in the reference (it scans every feature every round): exact
 * pruning of the fused round's search, on by default (LDS layout only).  A
 * feature is skipped when a lower bound on its error -- its last exact error
 * times the smallest sample-by-sample growth of the search weights since --
 * exceeds the smallest error already found in the round, so the chosen stump
 * is the full scan's.  pruning_stats gives the features scanned since the
 * stage began (a full scan is nFeatures per round). */
int vj_trainer_set_pruning(vj_trainer* t, int on);
int vj_trainer_pruning_stats(vj_trainer* t, int64_t* scanned);
int vj_trainer_set_schedule(vj_trainer* t, const int* sched);
int vj_trainer_schedule(vj_trainer* t, int* sched);

/* TileBestCandidate, include/vj/boosting_core.cuh (16 B): what
 * vj_trainer_select_best writes.  featureIdx is INT_MAX only when every
 * feature of the range is masked as used. */
typedef struct vj_tile_best {
    float err, theta;
    int32_t feature;
    int8_t parity, pad0, pad1, pad2;
} vj_tile_best;

/* ==== CPU legs of the compare verbs (host, single thread) ================
 * cuda_hello detect_compare / train_compare (src/main.cu:1032-1368) time the
 * same work on the CPU and on the GPU.  These are the CPU halves; no GPU
 * entry point uses them. */
/* detectMultiScaleCpu, src/main.cu:684-827: raw detections in (level, y, x)
 * order, stopping at maxDetections; maxScales > 0 limits the scale indices
 * visited.  *count receives the number found; min(*count, cap) are copied. */
int vj_cpu_detect_raw(const vj_feature* hFeatures, int nFeatures, const vj_stump* hStumps, int nStumps,
                      const vj_stage* hStages, int nStages, int winW, int winH, const uint8_t* hImage,
                      int w, int h, int stride, float scaleFactor, int minObjectSize, int maxDetections,
                      int maxScales, vj_detection* hOut, int cap, int* count);
/* The best stump over features [0, count) of n winW x winH patches
 * (positives labelled +1), as runCpuWeakSearch (src/main.cu:587-682) finds
 * it, with the search semantics of vj_trainer_find_threshold +
 * vj_trainer_select_best (exact fixed-point sums, GPU tie rules), so both
 * legs report the same stump.  hActive may be NULL (all active); weights
 * must be normalised (sum of |w| <= 1.5). */
int vj_cpu_best_weak(const uint8_t* hPatches, int n, int winW, int winH, const int8_t* hLabels,
                     const uint8_t* hActive, const float* hWeights, const vj_feature* hFeatures, int count,
                     vj_tile_best* out);

#ifdef __cplusplus
}
#endif
#endif /* VJ_API_H */
