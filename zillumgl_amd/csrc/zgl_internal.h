// zgl_internal.h -- launch interface between the C ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "zgl_device.h"

namespace zgl {

// Integrator identifiers (the reference's NaivePath / LightPath / TriplePath).
constexpr int IntegNaive = 0;
constexpr int IntegLight = 1;
constexpr int IntegTriple = 2;

struct PassParams {
    int firstPass;      // uSpp of the first pass in the launch
    int nPasses;        // passes rendered by this launch
    int freeCounter;    // uFreeCounter of the first pass (incremented per pass)
    int maxDepth;
    int russianRoulette;
    int sampleLight;
    int lightEnvUniform;
    float lightPortion;
    int pathsPerPass;   // light tracer: workgroup size 1536 x blocks per pass
    int loopsPerPass;   // triple LPT loops per thread
    float lptScale;     // triple LPT uScale
};

// Counter slots (uint64): 0 bytes, 1 rows, 2 leaf tests, 3 traversals, 4 samples
constexpr int NumCounters = 8;

// One scene of a batched persistent launch (device-resident table).
// Unit regions of the persistent schedules, one per XCD: each has its own counter (a
// batch has kUnitRegions counters per scene).
constexpr int kUnitRegions = 8;

struct BatchScene {
    DevScene s;
    PassParams prm;
    float4* frame;
    float4* res;    // prm.nPasses x filmW x filmH sample values
    int threshold;  // lanes still walking when a wave working on this scene stops to shade
};

hipError_t launchSobolTable(const uint32_t* sobolM, uint32_t* tab, int first, int count, hipStream_t st);
hipError_t launchFill(float* p, int64_t n, float v, hipStream_t st);
hipError_t launchSrgbExpand(const uint8_t* rgb, const float* lut, float4* out, int64_t n, hipStream_t st);
hipError_t launchPostProcess(const float4* frame, float4* out, unsigned char* rgb8, int w, int h, float scale, int tone,
                             hipStream_t st);
hipError_t launchFilmToFrame(const float* film, float4* frame, int w, int h, hipStream_t st);
hipError_t launchTraceClosest(const DevScene& s, const float* rays, int64_t n, int32_t* prim, float* t,
                              uint32_t* bin, unsigned long long* counters, hipStream_t st);
hipError_t launchTraceAny(const DevScene& s, const float* rays, const float* tmax, int64_t n,
                          uint8_t* occluded, hipStream_t st);
// res: per-sample values of up to resPasses passes (resPasses * W * H float4); the
// launch runs the passes in chunks of resPasses and adds each chunk to the frame in order.
hipError_t launchNaivePersistent(const DevScene& s, const PassParams& p, float4* frame, float4* res, int resPasses,
                                 int* pixelCounter, int threshold, int blocks, unsigned long long* counters,
                                 hipStream_t st);
// Workgroups of `threads` threads of `kernel` that one compute unit holds at once: the
// larger of `queried` (the runtime's occupancy query) and the residency computed from the
// kernel's VGPR count and static LDS for a CDNA3/4 compute unit (4 SIMDs, 512 VGPRs per
// lane allocated in granules of 8, at most 8 waves per SIMD, 160 KiB of LDS).  ROCm
// 7.2's runtime answered 2 for a 124-VGPR kernel of 256 threads that holds 4 per CU (so
// the ROCm 7.0 runtime, the measured launch times and the wave counters), and a process
// may bind the library to either runtime.  Launching more workgroups than fit is safe for
// these kernels (no workgroup waits on another); fewer leaves SIMDs idle.
int residentBlocksPerCU(const void* kernel, int threads, int queried);
int naivePersistentBlocksPerCU();
// The triple tracer's camera pass on the persistent schedule: p.nPasses passes in chunks
// of resPasses (res holds resPasses x W x H values), added to the planar film in order.
hipError_t launchTriplePersistent(const DevScene& s, const PassParams& p, float* film, float4* res, int resPasses,
                                  int* pixelCounter, int threshold, int blocks, hipStream_t st);
int triplePersistentBlocksPerCU();
// Traversal results logged per sample for the adjoint (see launchNaiveLogged).
constexpr int kHitLogQueries = 16;
// The persistent naive tracer over p's passes (all of them in one launch: res holds
// p.nPasses x W x H values, hitLog p.nPasses x W x H x kHitLogQueries entries), writing
// each sample's value to res and its traversal results to hitLog; the frame is untouched.
hipError_t launchNaiveLogged(const DevScene& s, const PassParams& p, float4* res, int2* hitLog, int* pixelCounter,
                             int threshold, int blocks, hipStream_t st);
int naiveBatchBlocksPerCU();
// maxPixels: the largest filmW x filmH of the batch (sizes the accumulate launch).
// The naive integrator's passes in FP64, one thread per pixel (zgl_reference.hip); sum
// holds W*H*3 per-pixel RGB sums and is added to.
hipError_t launchNaiveF64(const DevScene& s, const PassParams& p, double* sum, hipStream_t st);
hipError_t launchNaiveBatch(const BatchScene* items, int nScenes, int maxPixels, int* sceneCounters,
                            int blocks, hipStream_t st);
// Differentiable naive estimator: dFrame == null renders the replay estimator into
// frame; otherwise accumulates FP64 gradients into dMat (nMaterials x 5) and the envCopies
// copies of dEnv, summed into envSum (envW x envH x 3).
// hitLog (adjoint only, may be null): the traversal results of p's passes written by
// launchNaiveLogged; the replay then walks the MTBVH only for queries past the log.
hipError_t launchNaiveReplay(const DevScene& s, const PassParams& p, const float4* evalMat, const float4* evalEnv,
                             float4* frame, const float4* dFrame, double* dMat, double* dEnv, int envCopies,
                             double* envSum, const int2* hitLog, hipStream_t st);
// Global-queue path tracer: device bytes of the two item queues for a w x h film, blocks
// per CU of the extension kernel, and the launch of p.nPasses passes (counts: maxDepth + 1
// ints of device memory).
size_t globalQueueBytes(int w, int h);
int globalQueueBlocks();
hipError_t launchGlobalQueue(const DevScene& s, const PassParams& p, float4* frame, void* queues, int* counts,
                             int extendBlocks, hipStream_t st);
// Naive path tracer on the wavefront schedule (same frame as launchNaivePersistent):
// device bytes of its work buffer for a w x h film, blocks per CU of its trace and shade
// kernels, and the launch.
size_t wavefrontBytes(int w, int h, int passes);
// Passes rendered together (memory budget in bytes), spread evenly over the batches.
int wavefrontBatch(int w, int h, int nPasses, size_t budget);
void wavefrontBlocks(int& traceBlocksPerCU, int& shadeBlocksPerCU);
hipError_t launchNaiveWavefront(const DevScene& s, const PassParams& p, float4* frame, void* work, int batch,
                                int traceBlocks, int shadeBlocks, hipStream_t st);
hipError_t launchEnvExpand(const ushort4* env, float4* out, int64_t n, hipStream_t st);
// The light tracer (triple = false) or the triple tracer's light pass (triple = true)
// on the persistent schedule (zgl_lightpath.hip): p.nPasses passes of p.pathsPerPass
// units, splatted into the planar film; counter: one int of work queue.
hipError_t launchLightPersistent(const DevScene& s, const PassParams& p, bool triple, float* film, int* counter,
                                 int threshold, int blocks, hipStream_t st);
int lightPersistentBlocksPerCU(bool triple);
hipError_t launchLight(const DevScene& s, const PassParams& p, float* film, unsigned long long* counters,
                       hipStream_t st);
hipError_t launchTriplePT(const DevScene& s, const PassParams& p, float* film, unsigned long long* counters,
                          hipStream_t st);
hipError_t launchTripleLPT(const DevScene& s, const PassParams& p, float* film, unsigned long long* counters,
                           hipStream_t st);

}  // namespace zgl
