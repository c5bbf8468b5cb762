// Batched solves and the serving loop (filled in below).
#include "feec_internal.hpp"

using namespace feec;

extern "C" {

feec_status feec_batch_run(feec_ctx* ctx, int nreq, const feec_request* reqs,
                           const feec_solve_opts* opts, feec_result* results) {
  return guard([&] { fail(FEEC_E_UNSUPPORTED, "batch runner not built yet"); });
}
feec_status feec_server_create(feec_ctx* ctx, const feec_server_opts* opts, feec_server** out) {
  return guard([&] { fail(FEEC_E_UNSUPPORTED, "server not built yet"); });
}
feec_status feec_server_submit(feec_server* srv, const feec_request* req, int64_t* ticket) {
  return guard([&] { fail(FEEC_E_UNSUPPORTED, "server not built yet"); });
}
feec_status feec_server_poll(feec_server* srv, int64_t ticket, feec_result* result) {
  return guard([&] { fail(FEEC_E_UNSUPPORTED, "server not built yet"); });
}
feec_status feec_server_solution(feec_server* srv, int64_t ticket, double* u, int64_t* n) {
  return guard([&] { fail(FEEC_E_UNSUPPORTED, "server not built yet"); });
}
feec_status feec_server_drain(feec_server* srv) {
  return guard([&] { fail(FEEC_E_UNSUPPORTED, "server not built yet"); });
}
void feec_server_destroy(feec_server* srv) {}

}  // extern "C"
