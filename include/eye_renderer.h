/*
 * libEyeRenderer3 C ABI of the MI355X compound-eye renderer.
 *
 * The first block keeps the reference's exported symbols with identical names,
 * argument meaning and C signatures (libEyeRenderer3/libEyeRenderer.h:16-67),
 * so the reference's ctypes helpers
 * (python-examples/eyeRendererHelperFunctions.py:40-71) drive this library
 * unchanged.  Differences in behaviour, all additive:
 *   - headless: no window is created at load time, displayFrame is a no-op;
 *   - nothing throws across the ABI: a failing call records a message that
 *     eyeRendererGetLastError() returns and leaves the renderer usable;
 *   - errors the reference reports only on stderr keep their silent / no-op
 *     semantics (missing .eye file, compound setters on a regular camera).
 * The second block holds the additions: float readback, per-ommatidium sample
 * counts, random-stream control, batched rendering of many poses, ray-level
 * access, acceleration-structure inspection and the learned sampler.
 *
 * Plain C types only: pointers, sizes, floats.  The library owns every
 * returned pointer (valid until the next call that changes the same state).
 */
#ifndef EYE_RENDERER_MI355X_H
#define EYE_RENDERER_MI355X_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

/* libEyeRenderer3/libEyeRenderer.h:8-14 (32 bytes). */
typedef struct OmmatidiumPacket {
  float posX, posY, posZ;
  float dirX, dirY, dirZ;
  float acceptanceAngle;
  float focalpointOffset;
} OmmatidiumPacket;

/* Stand-in for CUDA's float3 in the two struct-returning calls: 12 bytes,
 * returned in registers exactly like the reference's float3
 * (eyeRendererHelperFunctions.py:6-9 declares it as three c_float). */
typedef struct EyeFloat3 {
  float x, y, z;
} EyeFloat3;

#ifdef __cplusplus
extern "C" {
#endif
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* ---- Configuration (libEyeRenderer.h:18-25, libEyeRenderer.cpp:211-295) ---- */
void setVerbosity(bool v);                  /* libEyeRenderer.h:19 */
void loadGlTFscene(const char* filepath);   /* libEyeRenderer.h:20, libEyeRenderer.cpp:215-220 */
void stop(void);                            /* libEyeRenderer.h:21, libEyeRenderer.cpp:289-295 */
void setRenderSize(int w, int h);           /* libEyeRenderer.h:22, libEyeRenderer.cpp:221-228 */
double renderFrame(void);                   /* libEyeRenderer.h:23, libEyeRenderer.cpp:229-242: host wall ms */
void displayFrame(void);                    /* libEyeRenderer.h:24: headless no-op */
void saveFrameAs(char* ppmFilename);        /* libEyeRenderer.h:25, libEyeRenderer.cpp:259-269: P6, flipped */
unsigned char* getFramePointer(void);       /* libEyeRenderer.h:26, libEyeRenderer.cpp:270-275: h*w*4 bytes */

/* ---- Camera control (libEyeRenderer.h:29-50, libEyeRenderer.cpp:306-388) ---- */
size_t getCameraCount(void);
void nextCamera(void);
void previousCamera(void);
size_t getCurrentCameraIndex(void);
const char* getCurrentCameraName(void);
void gotoCamera(int index);                 /* index wraps modulo the camera count (MulticamScene.cpp:928-933) */
bool gotoCameraByName(char* name);
void setCameraPosition(float x, float y, float z);
/* The reference takes float& x, y, z; C++ references are pointers in the ABI. */
void getCameraPosition(float* x, float* y, float* z);
void setCameraLocalSpace(float lxx, float lxy, float lxz, float lyx, float lyy, float lyz, float lzx,
                         float lzy, float lzz);
void rotateCameraAround(float angle, float axisX, float axisY, float axisZ);
void rotateCameraLocallyAround(float angle, float axisX, float axisY, float axisZ);
void translateCamera(float x, float y, float z);
void translateCameraLocally(float x, float y, float z);
void resetCameraPose(void);
/* Rotations about world X, then Y, then Z (radians), then translation
 * (libEyeRenderer.cpp:380-388). */
void setCameraPose(float posX, float posY, float posZ, float rotX, float rotY, float rotZ);

/* ---- Compound eye (libEyeRenderer.h:52-62, libEyeRenderer.cpp:393-468) ---- */
bool isCompoundEyeActive(void);
void setCurrentEyeSamplesPerOmmatidium(int s);        /* clamps to >= 1, restarts the random streams */
int getCurrentEyeSamplesPerOmmatidium(void);          /* -1 when the camera is not a compound eye */
void changeCurrentEyeSamplesPerOmmatidiumBy(int s);
size_t getCurrentEyeOmmatidialCount(void);            /* 0 when the camera is not a compound eye */
void setOmmatidia(OmmatidiumPacket* omms, size_t count);
const char* getCurrentEyeDataPath(void);
void setCurrentEyeShaderName(char* name);             /* projection: single_dimension_fast, single_dimension,
                                                         raw_ommatidial_samples */

/* ---- Scene queries (libEyeRenderer.h:64-66, MulticamScene.cpp:1757-1818) ---- */
bool isInsideHitGeometry(float x, float y, float z, char* name);
EyeFloat3 getGeometryMaxBounds(char* name);
EyeFloat3 getGeometryMinBounds(char* name);

/* libEyeRenderer.h:69 declares this with C++ linkage and returns the GLFW
 * window; headless here, so it is exported with C linkage and returns NULL. */
void* getWindowPointer(void);

/* ======================= Additions (no reference counterpart) =============== */

/* Message of the last failed call ("" when none); cleared by the next successful
 * loadGlTFscene or renderFrame. */
const char* eyeRendererGetLastError(void);

/* Linear (pre-gamma) RGB of every ommatidium of the last compound render,
 * 3 floats each; copies min(count, ommatidia) entries and returns that number. */
size_t getOmmatidialColours(float* rgb, size_t count);

/* Per-ommatidium sample counts for the current eye (count must equal the
 * ommatidial count; count == 0 returns to the uniform
 * setCurrentEyeSamplesPerOmmatidium value).  Counts of 0 render black. */
void setCurrentEyeSampleCounts(const uint32_t* counts, size_t count);
/* Total samples per frame of the current eye. */
uint64_t getCurrentEyeTotalSamples(void);

/* Random streams: key (seed, eye index), counter (ommatidium, sample, frame,
 * pose).  The frame index of the current eye advances after every render and
 * returns to 0 where the reference re-seeds (sample count or ommatidial count
 * changes). */
void setRandomSeed(uint32_t seed);
uint32_t getCurrentEyeFrameIndex(void);
void setCurrentEyeFrameIndex(uint32_t frame);
uint32_t getCurrentEyeIndex(void);   /* index among the scene's compound eyes, or 0xffffffff */

/* Pose of the current camera: 12 floats (position, local X, Y, Z axes), the
 * posed record the reference copies to the device (GenericCameraDataTypes.h:31-42).
 * Returns 0, or -1 when there is no camera. */
int getCameraPose(float* out12);

/* GPU milliseconds of the last renderFrame, measured with HIP events on the
 * render stream: out4 = (trace kernel, reduction kernel, projection kernel,
 * sum).  Returns 0. */
int getLastRenderTimings(double* out4);
/* Traversal work of the rays the next renderFrame() of the current eye will
 * trace: out3 = (BVH nodes visited, triangles tested, rays that hit).
 * Returns 0, or -1 on error. */
int getCurrentEyeTraversalStats(uint64_t* out3);

/* Rays the next renderFrame() of the current eye will trace, 8 floats each
 * (origin xyz, tmin, direction xyz, tmax), ray order (ommatidium, sample).
 * Returns the number written (at most `count`). */
size_t getCurrentEyeRays(float* rays, size_t count);
/* Per-sample radiance of the last render of the current eye, 4 floats each
 * (r, g, b, hit distance or -1 for a miss), same order. */
size_t getCurrentEyeSamples(float* samples, size_t count);

/* Closest hits of arbitrary rays (8 floats each as above) against the loaded
 * scene: primIds[i] = scene triangle index or -1, tbb[i*3..] = (t, b1, b2).
 * Returns 0 on success, -1 on error. */
int traceRays(const float* rays, size_t count, int32_t* primIds, float* tbb);

/* Failure-handling check of the wave-packet traversal (no reference
 * counterpart).  Its stack needs every lane of the wave: a packet traversal
 * entered by a partial wave flags the call, which then fails with an error,
 * and returns misses.  This traces rays64 (64 rays, 8 floats each) with one
 * wave whose lanes at or above activeLanes (1..64) skip the traversal.
 * ids64 receives each lane's hit id (-1 miss, -2 skipped).  Returns 1 when
 * the guard flagged the launch, 0 when not, -1 on error. */
int probePacketGuard(const float* rays64, int activeLanes, int32_t* ids64);

/* Batched rendering of the current eye at many poses in one launch sequence:
 * poses[6*k..] = (posX, posY, posZ, rotX, rotY, rotZ) with setCameraPose
 * semantics; rgbOut receives 3 floats per (pose, ommatidium), pose-major.
 * Pose k uses counter word 3 = k.  Returns GPU milliseconds, or -1 on error. */
double renderPoseBatch(const float* poses, size_t poseCount, float* rgbOut);

/* renderPoseBatch without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * devPoses: poseCount device-resident poses of 12 floats each, the layout of
 * getCameraPose (position, then the x, y and z axes; 4-byte aligned);
 * devColours: device output, 4 floats per (pose, ommatidium), pose-major:
 * (r, g, b, samples) (16-byte aligned).  The same units, random streams and
 * colours as renderPoseBatch at the same poses.  The work is enqueued on
 * `stream` (a hipStream_t; NULL: the library's stream) and not waited for.
 * The unit renderer's work counters live in a fixed 256-byte device array
 * (one per GPU, held until the process ends, cleared by a kernel at the start
 * of every call); nothing is allocated and no hipMalloc/hipFree is called.
 * Calls that overlap on one GPU race on that array.  Limits: at most
 * 2^32 - 65537 (pose, ommatidium) units per call; the current eye must be
 * resident, i.e. rendered by renderFrame or renderPoseBatch since its
 * ommatidia or sample counts last changed (those calls upload it and may
 * allocate).  A packet traversal fault is reported by the next call that
 * synchronises.  Returns 0, or -1 with the error set. */
int renderPoseBatchInto(const float* devPoses, size_t poseCount, float* devColours, void* stream);

/* ---- Learned sampler MLP (16 -> 64 -> 64 -> 64 -> 1, BF16 on MFMA) ----
 * Parameters are FP32 in the order W1[16][64], b1[64], W2[64][64], b2[64],
 * W3[64][64], b3[64], W4[64], b4.  Host arrays; all return 0 or -1. */
size_t mlpParamCount(void);
int mlpInit(uint32_t seed);
int mlpSetParams(const float* params, size_t count);
int mlpGetParams(float* params, size_t count);
int mlpPredict(const float* features, size_t rows, float* out);
/* Mean squared error over the rows and its gradient (FP32, mlpParamCount). */
int mlpLossGrad(const float* features, const float* targets, size_t rows, float* loss, float* grad);
/* Adam minibatch training; epochLosses (nullable) receives one mean loss per epoch. */
int mlpTrain(const float* features, const float* targets, size_t rows, int epochs, int batch, float lr,
             uint32_t seed, float* epochLosses);

/* ---- Learned adaptive sampling of the current eye ----
 * Training rows for the sampler: for every (pose, ommatidium) with
 * unitMask[pose * ommatidia + omm] != 0 (NULL mask: all), 16 features from
 * `pilot` samples (the unit's own statistics and those of its 6 nearest
 * ommatidia, nearestOmmatidia; the pilot covers every unit of the poses) and
 * the label log2(max(1, mean RGB variance / errorTarget^2)) from `labelSpp`
 * further samples.  Writes at most maxRows rows (features 16 floats each);
 * returns the number of rows, or -1. */
int64_t adaptiveTrainingSet(const float* poses, size_t poseCount, const uint8_t* unitMask, int pilot, int labelSpp,
                            float errorTarget, float* features, float* labels, size_t maxRows);
/* Neighbour table of the learned sampler (no reference counterpart; host
 * computation, no GPU): out[i * k + m] = the m-th nearest of the other `count`
 * ommatidia (8 floats each, the OmmatidiumPacket layout) to ommatidium i under
 * d^2 = sum over the unit axis and position / 0.01 of squared differences
 * (FP64, in that coordinate order), ties to the lower index; slots beyond
 * count - 1 neighbours repeat i.  k in 1..16; returns 0, or -1. */
int nearestOmmatidia(const float* ommatidia, size_t count, int k, uint32_t* out);
/* The two stages of the last renderPoseBatchAdaptive, per (pose, ommatidium)
 * unit: pilot and top-up colours as (r, g, b, samples) and the top-up sample
 * counts (any pointer may be NULL).  Returns maxUnits, or -1 when the last
 * render had fewer units. */
int64_t getAdaptiveStages(float* pilot, float* topup, uint32_t* extra, size_t maxUnits);
/* Error the sampler network's labels were made for (default 0.01). */
int setAdaptiveLabelError(float labelError);
/* Measurement switch (no reference counterpart): while on, every wave-packet
 * traversal of renderFrame, renderPoseBatch and renderPoseBatchAdaptive adds
 * its node visits and the triangles of the leaves it enters, counted once
 * per wave (what the wave fetches), to a device counter pair.  Off by
 * default; the counting adds one atomic per wave. */
void setPacketCounting(bool enabled);
/* Reads the counter pair into out2 = (node visits, leaf triangles) and clears
 * it.  Returns 0, or -1. */
int getPacketCounts(uint64_t* out2);
/* Where renderPoseBatchAdaptive gets each unit's required samples: 0 (the
 * default) the trained network; 1 the network-free estimate from the pooled
 * pilot standard deviation of the unit and its 6 nearest ommatidia, sigma =
 * (own + 6 * neighbours' mean) / 7, required samples sigma^2 / labelError^2.
 * Everything after that step is shared.  Returns 0, or -1 for another mode. */
int setAdaptiveSampler(int mode);
/* Device milliseconds of the last renderPoseBatchAdaptive's stages, from HIP
 * events on the library's stream: out4 = (pilot render, features + required
 * samples + allocation, top-up render, combination).  Returns 0, or -1. */
int getAdaptiveTimings(double* out4);
/* Batched adaptive render: 8-sample pilot, MLP prediction, per-pose Neyman
 * allocation for RMS error errorTarget (times budgetScale), top-up, combine.
 * The network predicts the samples needed for the label error (the
 * errorTarget of the last adaptiveTrainingSet call, or setAdaptiveLabelError);
 * allocations scale by (label error / errorTarget)^2.
 * rgbOut / countsOut (nullable) per (pose, ommatidium); *totalSamples gets the
 * samples traced.  Returns GPU milliseconds, or -1. */
double renderPoseBatchAdaptive(const float* poses, size_t poseCount, float errorTarget, int maxSpp,
                               float budgetScale, float* rgbOut, uint32_t* countsOut, uint64_t* totalSamples);

/* Acceleration-structure inspection: info[0..7] = triangle count, node count,
 * stack bound, levels, leaf size, build ms * 1000, bytes per node (128 per
 * four child slots), 0.  Returns 0 or -1. */
int getSceneBVHInfo(int64_t* info);
/* Copies the 4-wide node array (128 bytes per node); returns bytes copied. */
size_t copySceneBVHNodes(void* dst, size_t bytes);
/* Copies sorted triangle ids and Morton codes (n entries each). */
size_t copySceneSortOrder(uint32_t* ids, uint32_t* codes, size_t count);
/* Rebuilds the BVH of the loaded scene with another leaf size (1..16). */
int setSceneLeafSize(int leafSize);

/* Random-stream index of the first sample of every ommatidium in renderFrame
 * and renderPoseBatch (default 0): samples base .. base + n - 1 are drawn, so
 * a later render can add samples independent of an earlier one. */
void setCurrentEyeSampleBase(uint32_t base);
/* Traversal strategy of eye rays: 0 = one ray per lane (own LDS stack),
 * 1 = wave packets, 2 = wave packets reading nodes and triangles through the
 * scalar unit (default).  Results are identical; only speed differs.
 * autotuneRenderBatch sets it to the tuned winner. */
void setTraversalMode(int mode);
/* Keep the per-sample records of renderFrame for getCurrentEyeSamples (off by
 * default: the packet renderer reduces samples in registers). */
void setSampleRecording(bool enabled);
/* Kernel of the last compound-eye render: "k_render_units", or for a split
 * launch (uniform counts and few units, where each unit's 64-sample chunks
 * run on different waves) "k_render_tasks" (wave packets, any count) or
 * "k_render_chunks" (per-ray traversal, above 32 samples). */
const char* getLastRenderKernel(void);
/* Allows split launches (default true); results are the same either way. */
void setSplitLaunches(bool enabled);
/* Long-task split of the wave-packet task kernel (an eye rendered frame after
 * frame at a uniform rate): the tasks (64-sample chunks) that took the most
 * packet steps in the previous render of the same eye are traced first, the
 * longest as four or two packets of neighbouring-elevation rays, the next
 * whole.  mode 0: never; 1: always;
 * 2 (default): per eye, by the measured trace times of a few renders with and
 * without it, tried again every 1,000 renders and after any change of the
 * rays.  Results are the same either way.  No counterpart in the reference
 * (scheduling of this renderer). */
void setTaskSplit(int mode);
/* Tasks of the last renderFrame of the current eye traced as several packets
 * (0 when the split was off for that render or nothing was long enough).  No counterpart
 * in the reference (diagnostic of this renderer's scheduling). */
int getCurrentEyeTaskSplits(void);
/* Launch shape of the persistent unit renderer: ommatidia taken from the work
 * queue per grab (0 = automatic, at most 64: a wave holds one grab's sample
 * counts in its lanes), and blocks of 256 threads (0 = every resident slot
 * once).  The adaptive passes keep the blocks and traversal but choose their
 * own grab size. */
void setRenderTuning(int unitsPerGrab, int blocks);
/* Writes the current launch shape: units per grab, blocks, traversal mode. */
int getRenderTuning(int* shape3);
/* Launch-shape autotuning of renderPoseBatch (no reference counterpart; the
 * reference has a single launch shape).  When enabled, the first batch of each
 * workload (scene, eye, sample counts) times the candidate shapes on a sample
 * of its poses and later batches reuse the winner. */
void setBatchAutotune(int enabled);
/* Merges the tuned shapes of a text file with those already known, writes
 * them all back and saves later tunes there (NULL: stop saving); returns the
 * number of cached entries, -1 on a malformed or unwritable file. */
int setAutotuneCacheFile(const char* path);
/* Times every candidate shape on (a sample of) these poses and installs the
 * fastest as the shape of all later unit launches.  trials receives rows of
 * 4 floats (units per grab, blocks, traversal: 0 packet, 1 packet with scalar
 * loads, 2 per-ray; ms); returns the number of candidates, -1 on error. */
int autotuneRenderBatch(const float* poses, size_t poseCount, int repeats, float* trials, size_t maxTrials);

/* Bandwidth reference for profiling: copies `bytes` (rounded down to 16-byte
 * units) between two device buffers with a plain vector copy kernel
 * (k_copy_probe), `repeats` times; returns the mean GPU ms per copy or -1. */
double runCopyProbe(size_t bytes, int repeats);

/* Primary rays (8 floats: origin, tmin, direction, tmax) of every pixel of the
 * current regular camera at its pose and the render size, row-major; returns
 * the number written (maxRays = 0: the number of pixels), 0 for a compound eye. */
size_t getCurrentCameraRays(float* rays, size_t maxRays);

/* Host-only glTF parse (no GPU): world-space triangles (9 floats), colours
 * (9 floats) and colour modes of `path`.  Returns the triangle count (call
 * with maxTris = 0 to query), or -1 on error. */
int64_t parseGlTFTriangles(const char* path, float* pos, float* col, uint32_t* mode, size_t maxTris);
/* Host-only parse of the base-colour textures of `path`: 6 ints per texture
 * (width, height, wrap s, wrap t, filter, first texel; wrap 0 repeat, 1 clamp,
 * 2 mirrored repeat; filter 0 point, 1 linear) and the decoded RGBA8 texels
 * (red in the low byte).  A textured triangle's mode is 2 | texture << 2 and
 * its colour slots hold u0 v0 u1 v1 u2 v2.  Returns the texture count and
 * sets *texelCount, or -1 on error. */
int64_t parseGlTFTextures(const char* path, int32_t* desc, size_t maxTextures, uint32_t* texels, size_t maxTexels,
                          int64_t* texelCount);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif

#endif /* EYE_RENDERER_MI355X_H */
