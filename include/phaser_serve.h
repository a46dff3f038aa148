/*
 * phaser_serve.h -- a request loop over the phaser_amd C ABI: scan pairs of
 * mixed bandwidth and voxel count arrive at any time and are registered in
 * batches on one GPU.
 *
 * No counterpart in the reference, which registers one pair per call
 * (BaseRegistration::registerPointCloud, base-registration.h:16-25).  Each
 * request is exactly one such call; its result equals what phz_register
 * returns for the same pair on a context of the same shape.
 *
 * Model.  A *shape* is a (bandwidth, n_voxels) pair.  Per shape the server
 * keeps up to contexts_per_shape phaser_amd contexts (one HIP stream and one
 * worker thread each, tables resident).  A worker takes every queued request
 * of its shape, up to max_batch, as soon as it is free (continuous batching:
 * requests that arrive while a batch runs join the next one), so batches of
 * different shapes -- and of one shape on different contexts -- overlap on
 * the device.  Requests with the same nonzero target_key and shape are
 * batched around one target, whose spectra are then computed once.
 * Contexts are admitted against hbm_budget with phz_estimate_memory: a shape
 * whose first context does not fit waits until idle contexts of other shapes
 * are released (least recently used first).
 *
 * Point buffers are read by the worker: they must stay valid until the
 * request's response has been returned by phs_wait or phs_poll.
 * All functions are thread-safe.  Status codes as in phaser_amd.h.
 */
#ifndef PHASER_SERVE_H_
#define PHASER_SERVE_H_

#include <stddef.h>
#include <stdint.h>

#include "phaser_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phs_server phs_server;

typedef struct {
  int32_t device;
  int32_t max_batch;           /* pairs per batch and context */
  int32_t contexts_per_shape;  /* >= 1 */
  int32_t discretize_lower;    /* voxel grid bounds of every shape */
  int32_t discretize_upper;
  int32_t autotune;            /* nonzero: phz_autotune each new context */
  int64_t max_points;          /* per cloud */
  uint64_t hbm_budget;         /* device bytes all contexts together may hold */
  double batch_window_ms;      /* a free worker with fewer than max_batch
                                  requests waits this long for more (0: none) */
  int32_t precision;           /* phz_config.precision of every context */
  int32_t graphs;              /* nonzero: contexts replay their batches from HIP graphs
                                  (phz_set_graphs); default 1 */
  int32_t bingham_peak_neighbors;   /* phz_config fields of every context */
  int32_t gaussian_peak_neighbors;
  int32_t method;                   /* phz_config.method: PHZ_METHOD_SPH_OPT or PHZ_METHOD_SPH */
  int32_t skip_rot_axes;            /* phz_config.skip_rot_axes (PHZ_METHOD_SPH) */
} phs_config;

typedef struct {
  int32_t bandwidth;
  int32_t n_voxels;
  uint64_t target_key;  /* 0: independent target; else shared by equal keys */
  const float* target;  /* cloud_prev, Pt x (x, y, z, intensity) */
  int64_t Pt;
  const float* source;  /* cloud_cur, Ps x 4 */
  int64_t Ps;
  uint64_t tag;         /* returned unchanged */
} phs_request;

typedef struct {
  uint64_t ticket;
  uint64_t tag;
  int32_t status;         /* 0, or the phaser_amd status of the failed batch */
  int32_t batch_pairs;    /* pairs in the batch that served the request */
  int32_t shared_target;  /* the batch was registered against one target */
  int32_t reserved;
  phz_result result;
  double queue_ms;        /* submit -> batch start */
  double service_ms;      /* batch start -> batch end */
  double latency_ms;      /* submit -> batch end */
  char error[160];
} phs_response;

typedef struct {
  uint64_t submitted, completed, failed, batches, pairs_in_batches;
  uint64_t contexts_created, contexts_released;
  uint64_t bytes_in_use, bytes_peak;
  int32_t contexts_live;
  int32_t queued;
} phs_stats;

/* max_batch 16, contexts_per_shape 2, bounds -50/50, max_points 262144,
 * hbm_budget 64 GiB, batch_window_ms 0. */
void phs_default_config(phs_config* cfg);
int phs_create(const phs_config* cfg, phs_server** out);
/* Finishes every queued request, then releases all contexts. */
void phs_destroy(phs_server* srv);
/* Queues one request; fails at once (-1/-3) for a shape that cannot be
 * built, a cloud above max_points, or a context larger than hbm_budget. */
int phs_submit(phs_server* srv, const phs_request* req, uint64_t* ticket);
/* Blocks until the request is done and returns its response (once). */
int phs_wait(phs_server* srv, uint64_t ticket, phs_response* out);
/* Returns up to max finished responses not yet returned, waiting up to
 * wait_ms for the first one (0: no wait, negative: no limit). */
int phs_poll(phs_server* srv, phs_response* out, int max, int* count, int wait_ms);
/* Blocks until every submitted request has finished. */
int phs_drain(phs_server* srv);
int phs_get_stats(phs_server* srv, phs_stats* out);
const char* phs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* PHASER_SERVE_H_ */
