// dachshund.h -- drop-in of lib/dachshund.h (umbrella header), plus the
// GPU context used by the drop-in implementation.
#ifndef DS_DROPIN_DACHSHUND_H
#define DS_DROPIN_DACHSHUND_H

#include "linalg.h"
#include "map.h"
#include "map_covar.h"
#include "pcg.h"
#include "pixel.h"
#include "signal_covar.h"
#include "timer.h"
#include "util.h"
#include "wf_vector.h"

struct ds_ctx;
// The process-wide GPU context of the drop-in classes: device
// $DACHSHUND_DEVICE (default 0), created on first use.  Aborts with a
// message if no GPU or no libdachshund_gpu.so is available: there is no
// host fallback.
ds_ctx* dachshund_gpu_context();

#endif
