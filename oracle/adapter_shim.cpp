// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// The backend entry points of include/banjo_lattice.h implemented over the
// oracle restatement (build/liboracle.so, orc_* exports), so that the
// reference-side adapter (banjo_amd/adapter/hip_lattice_backend.cpp) can be
// driven end to end on a machine with no GPU: the adapter is linked against
// this file instead of the product library, and its answers are compared with
// the reference's own CPU backend in adapter_check.cpp. The oracle equals the
// reference's lattice lane bit for bit (tests/test_oracle.py), so any field the
// adapter copies wrongly shows up as a difference. Only the check program links
// this; nothing in the product does.

#include "../include/banjo_lattice.h"

#include <cstring>

extern "C" {
// The oracle's exports (oracle/lattice_oracle.cpp).
const char *orc_last_error(void);
int orc_backend_create(int precision, const blx_schedule_view *sv, void **out);
int orc_backend_destroy(void *h);
int orc_backend_upload(void *h, const blx_state_view *st, const blx_settings *s, const blx_sphere *sp);
int orc_backend_set_external_forces(void *h, const double *f, uint64_t substeps);
int orc_backend_run(void *h, const blx_run_control *c, blx_run_status *st);
int orc_backend_status(void *h, blx_run_status *st);
int orc_backend_download(void *h, blx_state_view *st, blx_sphere *sp);
int orc_backend_first_failure_bonds(void *h, uint32_t *out, uint64_t cap, uint64_t *count);
int orc_backend_frame_count(void *h, uint32_t *count);
int orc_backend_take_frame(void *h, uint32_t index, uint64_t *step, float *u, uint8_t *alive, float *damage,
                           blx_sphere *sp);
int orc_backend_clear_frames(void *h);

#define ORC(b) reinterpret_cast<void *>(b)

const char *blx_last_error(void) { return orc_last_error(); }
int blx_available(void) { return 1; }
int blx_description(char *out, size_t cap) {
    if (out && cap) {
        std::strncpy(out, "oracle restatement (CPU)", cap - 1);
        out[cap - 1] = 0;
    }
    return BLX_OK;
}
uint32_t blx_multiprocessor_count(void) { return 1; }
int blx_backend_create(int /*device*/, int precision, const blx_schedule_view *sv, blx_backend **out) {
    void *h = nullptr;
    const int rc = orc_backend_create(precision, sv, &h);
    *out = reinterpret_cast<blx_backend *>(h);
    return rc;
}
int blx_backend_destroy(blx_backend *b) { return orc_backend_destroy(ORC(b)); }
int blx_backend_name(const blx_backend *, char *out, size_t cap) { return blx_description(out, cap); }
int blx_backend_upload(blx_backend *b, const blx_state_view *st, const blx_settings *s, const blx_sphere *sp) {
    return orc_backend_upload(ORC(b), st, s, sp);
}
int blx_backend_set_external_forces(blx_backend *b, const double *f, uint64_t n) {
    return orc_backend_set_external_forces(ORC(b), f, n);
}
int blx_backend_run(blx_backend *b, const blx_run_control *c, blx_run_status *st) { return orc_backend_run(ORC(b), c, st); }
int blx_backend_status(const blx_backend *b, blx_run_status *st) {
    return orc_backend_status(ORC(const_cast<blx_backend *>(b)), st);
}
int blx_backend_download(blx_backend *b, blx_state_view *st, blx_sphere *sp) { return orc_backend_download(ORC(b), st, sp); }
int blx_backend_first_failure_bonds(const blx_backend *b, uint32_t *out, uint64_t cap, uint64_t *count) {
    return orc_backend_first_failure_bonds(ORC(const_cast<blx_backend *>(b)), out, cap, count);
}
int blx_backend_frame_count(const blx_backend *b, uint32_t *n) {
    return orc_backend_frame_count(ORC(const_cast<blx_backend *>(b)), n);
}
int blx_backend_take_frame(blx_backend *b, uint32_t k, uint64_t *step, float *u, uint8_t *alive, float *damage,
                           blx_sphere *sp) {
    return orc_backend_take_frame(ORC(b), k, step, u, alive, damage, sp);
}
int blx_backend_clear_frames(blx_backend *b) { return orc_backend_clear_frames(ORC(b)); }
}
