// Serving: a stream of SAT / QSAT / SATOSTOP requests solved by worker lanes that run
// concurrently on one device.
#pragma once
#include <vector>

#include "../../include/naivesat_hip.h"
#include "circuit.h"
#include "runtime.h"

namespace ns {

struct ServeRequest {
  uint32_t kind = 0;  // NS_REQ_SAT / NS_REQ_QSAT / NS_REQ_SATOSTOP
  uint32_t unknowns = 0;
  const Circuit* circuit = nullptr;
  std::vector<uint8_t> quants;
  double arrival_s = 0.0;
};

// Device bytes a request needs besides the lane's small buffers (admission control).
uint64_t request_hbm_bytes(const ServeRequest& r);

void serve(Context& ctx, const std::vector<ServeRequest>& reqs, const ns_serve_opts& o,
           std::vector<ns_request_result>& out, double* total_s);

}  // namespace ns
