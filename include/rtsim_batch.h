/*
 * rtsim_batch.h -- batched pings of the rtsim C ABI.
 *
 * The reference computes one ping per SensorInstance::compute_output call
 * (SensorInstance.h:67) and synchronises the device after every kernel
 * (Binner.cu:78,118,134, SensorInstance2D.cu:100,117).  A batch computes many
 * sensor instances (mixed 1D / 2D, uneven sizes, per-sensor form) together:
 * each stage of the path (stream, the four sort passes, segments, windows,
 * the convolutions) is ONE launch covering the workgroups of every ping, and
 * the ~20 launches are captured into one hipGraph on the first compute and
 * again only after a sensor's ranges, sample count or tile shape change.
 */
#ifndef RTSIM_BATCH_H
#define RTSIM_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "rtsim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtsim_batch_s* rtsim_batch;

/* forms[i] is the RTSIM_FORM_* of sensors[i] (NULL: plain FP32 for all).
 * Form-specific device state is prepared here, so an unsupported form is
 * reported at creation.  The sensors must outlive the batch.                  */
int rtsim_batch_create(const rtsim_sensor* sensors, const int* forms, uint32_t count,
                       rtsim_batch* out);
/* Includes the streaming kernel (rtsim_receiver_write_samples) of sensor
 * `index` in the batch: n must equal its sample count.  hits == NULL unbinds.
 * Binding the same arrays again keeps the captured graph.                     */
int rtsim_batch_bind_hits(rtsim_batch batch, uint32_t index, const rtsim_hits* hits, size_t n);
/* Queues the whole batch on `stream` (one graph launch).  A receiver pose set
 * since the last compute (rtsim_sensor_set_pose) is copied into that ping's
 * streaming arguments on `stream` before the launch, without re-capture; the
 * graph is re-captured after a sensor's ranges, sample count, tile shape or
 * bound hit arrays change.                                                    */
int rtsim_batch_compute(rtsim_batch batch, void* stream);
/* Tile and launch-mode autotuner of the batch (no counterpart in the
 * reference, whose pings are computed one at a time): for each class of
 * MFMA-form sensors (same grid and PSF size) in turn, times `iterations`
 * graph replays of the whole batch with every correlation tile shape and wave
 * count that fits the class, keeps the fastest (the other classes held at
 * their current choice) and sets it on those sensors
 * (rtsim_sensor_set_tile_shape / _waves).  Then does the same in the other
 * launch mode (rtsim_batch_set_persistent) and keeps the faster mode with its
 * tiles.  A full batch can prefer other tiles than one ping alone
 * (rtsim_sensor_autotune).  Leaves every ping computed with the chosen mode
 * and shapes; best_ms (may be NULL) receives the batch's mean replay time.
 * Synchronises the stream.                                                     */
int rtsim_batch_autotune(rtsim_batch batch, int iterations, void* stream, float* best_ms);
/* Persistent mode (no counterpart in the reference): on != 0 runs every
 * grid-form (MFMA) ping of the batch in ONE persistent launch instead of the
 * per-stage launches: workgroups take the tasks of every stage of every
 * grid-form ping (streaming tiles, bucket rows, tiles, buckets, output tiles)
 * in level-major order from an atomic ticket, each task waiting only for its
 * own ping's previous stage (a correlation tile only for the deposit buckets
 * its rows read), so the stages of different pings overlap.  The launch is
 * captured in the batch's hipGraph like the others; plain-form pings keep
 * their per-stage launches.  It runs 8-wave correlation tiles: turning it on
 * sets 8 waves on every grid-form sensor of the batch (their current tile
 * shape when it allows 8 waves, else the first shape that does;
 * RTSIM_ERR_UNSUPPORTED when none does).  Pings are bit-identical to the
 * sensors' own computes.  Takes effect at the next compute (re-capture).      */
int rtsim_batch_set_persistent(rtsim_batch batch, int on);
int rtsim_batch_persistent(rtsim_batch batch, int* on);
/* Waits for the batch's last launch and reports whether its persistent
 * launch completed: every wait of that launch is bounded (two seconds); a
 * wait that exceeds the bound abandons the launch (its pings are then
 * undefined) and this returns RTSIM_ERR_HIP.  RTSIM_OK otherwise, also
 * without a persistent launch.                                                */
int rtsim_batch_check(rtsim_batch batch);
/* Waits for the batch's last launch, then frees it.                            */
int rtsim_batch_destroy(rtsim_batch batch);

#ifdef __cplusplus
}
#endif

#endif /* RTSIM_BATCH_H */
