// batch.hpp -- many transforms per call: lanes of HIP streams, graph capture,
// lane-count tuning and completion tickets (entry points in so3x.h).
#pragma once
#include "gpu_common.hpp"
#include "so3/so3x.h"

namespace so3g {

// mode: SO3X_BATCH_AUTO (the tuned mode and lane count for this batch's
// shape when lanes <= 0, else lanes), SO3X_BATCH_LANES or
// SO3X_BATCH_PERSISTENT
void batch_enqueue(const so3x_item_t *items, int count, int lanes, int mode, hipStream_t st);
int batch_tune(const so3x_item_t *items, int count, hipStream_t st);
void batch_choice(const so3x_item_t *items, int count, int *mode, int *lanes);
long long graph_create(const so3x_item_t *items, int count, int lanes, int mode);
void graph_launch(long long graph, hipStream_t st);
void graph_destroy(long long graph);
long long ticket_new(hipStream_t st);
int ticket_done(long long ticket);
void ticket_wait(long long ticket);
void release_batch();

}  // namespace so3g
