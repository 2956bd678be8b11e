// batch.cpp — many independent instances (own width, circuit, observable).
//
// Stands in for the sequential parameter-set loops of batch_expectation
// (src/qarpx/src/simulator/qarp_simulator.cpp:971-984) and the per-point
// sweep of QarpEngine._sweep (src/qarp/engines/_qarp_engine.py:450-503).
// Instances are grouped by register width so one device state per width is
// reused; each instance runs the fused forward apply, H|psi>, and (when
// gradients are requested) the adjoint reverse sweep.
#include "common.h"

#include <map>

namespace qhip {

void batch_energy_grad(const qhip_config* cfg, int n_instances, const int32_t* n_qubits,
                       const qhip_stream* streams, const qhip_observable* observables,
                       const int32_t* n_params, int precision, double* energies, double* grads,
                       const int64_t* grad_offsets) {
    if (precision != QHIP_C128 && precision != QHIP_C64) fail(QHIP_INVALID, "batch: bad precision");
    std::map<int, std::vector<int>> by_width;
    for (int i = 0; i < n_instances; ++i) {
        if (n_qubits[i] < 0 || n_qubits[i] > 34) fail(QHIP_INVALID, "batch: n_qubits out of range");
        by_width[n_qubits[i]].push_back(i);
    }
    for (const auto& [n, members] : by_width) {
        qhip_state st;
        st.n = n;
        st.precision = precision;
        if (hipMalloc(&st.d, st.dim() * st.amp_bytes()) != hipSuccess) {
            (void)hipGetLastError();
            fail(QHIP_NOMEM, "batch: state allocation failed");
        }
        try {
            for (int i : members) {
                state_reset(&st);
                if (grads) {
                    run_gradient(cfg, &streams[i], n_params[i], &st, &observables[i], nullptr,
                                 grads + grad_offsets[i], &energies[i]);
                } else {
                    apply_stream(cfg, &streams[i], &st, nullptr, false);
                    energies[i] = transition(&st, &st, &observables[i]).real();
                }
            }
        } catch (...) {
            (void)hipStreamSynchronize(ctx().stream);
            (void)hipFree(st.d);
            throw;
        }
        QHIP_CHECK(hipStreamSynchronize(ctx().stream));
        QHIP_CHECK(hipFree(st.d));
    }
}

}  // namespace qhip
