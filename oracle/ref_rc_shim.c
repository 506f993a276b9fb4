/* ref_rc_shim.c - test harness around the reference's rate controller
 * (approach1-compute-encoder/src/rate_control.c, built from the reference
 * tree into _ref/libref_rc.so): allocation and the fields a test sets or
 * reads, so that ctypes needs no copy of the struct layout. */
#include <stdlib.h>

#include "rate_control.h"

rate_control_t *ref_rc_new(void) { return (rate_control_t *)calloc(1, sizeof(rate_control_t)); }
void ref_rc_free(rate_control_t *rc) { free(rc); }
void ref_rc_set_model(rate_control_t *rc, int on) { rc->model = on; }
void ref_rc_set_current_qp(rate_control_t *rc, int qp) { rc->current_qp = qp; }
long long ref_rc_fullness(const rate_control_t *rc) { return (long long)rc->buffer_fullness; }
long long ref_rc_integral(const rate_control_t *rc) { return (long long)rc->error_integral; }
double ref_rc_debt(const rate_control_t *rc) { return rc->debt; }
int ref_rc_base_qp(const rate_control_t *rc) { return rc->base_qp; }
unsigned ref_rc_bits_per_frame(const rate_control_t *rc) { return rc->target_bits_per_frame; }
