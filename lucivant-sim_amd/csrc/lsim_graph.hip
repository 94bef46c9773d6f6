// Library-owned HIP graphs: a caller records a sequence of this library's launches on one stream
// once and replays it with one launch. Capture goes through HIP's stream capture in thread-local
// mode, so only calls from the capturing thread are recorded (and checked): the entry points of
// this library only enqueue kernels on the stream they are given, which is what a graph holds.
// A captured graph replays with the same device pointers and kernel arguments; data it reads
// (matrices, states) may change between replays.
#include <new>

#include "lsim_internal.h"

using namespace lsim;

struct lsim_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t nodes = 0;
};

extern "C" {

int lsim_graph_capture_begin(void* stream) {
  LSIM_REQUIRE(stream != nullptr, "graph capture needs a stream of its own (not the null stream)");
  hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(as_stream(stream), &status) != hipSuccess)
    return fail(LSIM_EHIP, "hipStreamIsCapturing failed");
  LSIM_REQUIRE(status == hipStreamCaptureStatusNone, "the stream is already capturing");
  const hipError_t e = hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail(LSIM_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
  return LSIM_OK;
}

int lsim_graph_capture_end(void* stream, lsim_graph** out) {
  LSIM_REQUIRE(stream != nullptr && out != nullptr, "the stream or the output pointer is NULL");
  *out = nullptr;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(as_stream(stream), &graph);
  if (e != hipSuccess || graph == nullptr)
    return fail(LSIM_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  size_t nodes = 0;
  if (hipGraphGetNodes(graph, nullptr, &nodes) != hipSuccess) nodes = 0;
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(LSIM_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  lsim_graph* g = new (std::nothrow) lsim_graph;
  if (g == nullptr) {
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    return fail(LSIM_EINVAL, "out of host memory for the graph handle");
  }
  g->graph = graph;
  g->exec = exec;
  g->nodes = nodes;
  *out = g;
  return LSIM_OK;
}

long long lsim_graph_nodes(const lsim_graph* g) { return g == nullptr ? -1 : (long long)g->nodes; }

int lsim_graph_launch(lsim_graph* g, void* stream) {
  LSIM_REQUIRE(g != nullptr && g->exec != nullptr, "the graph handle is NULL");
  const hipError_t e = hipGraphLaunch(g->exec, as_stream(stream));
  if (e != hipSuccess) return fail(LSIM_EHIP, "hipGraphLaunch: %s", hipGetErrorString(e));
  return LSIM_OK;
}

int lsim_graph_destroy(lsim_graph* g) {
  if (g == nullptr) return LSIM_OK;
  hipError_t e1 = g->exec ? hipGraphExecDestroy(g->exec) : hipSuccess;
  hipError_t e2 = g->graph ? hipGraphDestroy(g->graph) : hipSuccess;
  delete g;
  if (e1 != hipSuccess || e2 != hipSuccess) return fail(LSIM_EHIP, "destroying the graph failed");
  return LSIM_OK;
}

}  // extern "C"
