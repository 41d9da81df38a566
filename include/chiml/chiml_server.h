/*
 * chiml_server.h -- C ABI of the sweep server (continuous batching on one GPU).
 *
 * The reference runs one simulation per MPI job (src/main.cpp:40-118); a
 * parameter sweep is many such jobs.  This server keeps many simulations
 * device-resident at once: requests are admitted while their estimated HBM
 * footprint fits the budget and a slot is free, each admitted simulation is
 * built on a builder thread, then stepped in chunks on its own HIP stream
 * (optionally as captured HIP graphs) by a stepping thread that polls every
 * stream and refills whichever has finished its chunk, so the work of
 * different simulations overlaps on the device.  Requests that arrive while
 * others run are admitted as soon as memory and slots free up.
 *
 * All calls return 0 or a negative CHIML_E_* code (chiml_dev.h); the message
 * is in chiml_sim_last_error() of the calling thread.  The server object may
 * be used from several host threads.
 */
#ifndef CHIML_SERVER_H
#define CHIML_SERVER_H

#include <stddef.h>

#include "chiml_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chiml_server chiml_server_t;

enum { CHIML_REQ_QUEUED = 0, CHIML_REQ_BUILDING = 1, CHIML_REQ_RUNNING = 2, CHIML_REQ_DONE = 3, CHIML_REQ_FAILED = 4 };

/* hbm_budget: device bytes the server may hold for simulations (admission
 * uses chiml_sim_estimate_bytes); max_active: simulations admitted at once;
 * chunk_steps: steps enqueued per simulation per refill; graph_steps: steps
 * per captured HIP graph (0 = plain launches). */
int chiml_server_create(size_t hbm_budget, int max_active, int chunk_steps, int graph_steps, chiml_server_t** out);
/* enable = 1: the simulations whose last chunk has finished take their next
 * chunk together, through chiml_sim_step_batch (one launch per kernel class
 * and phase for all of them; per-simulation graphs are then not used);
 * 0 (the default): each simulation enqueues its chunks on its own stream. */
int chiml_server_set_batch(chiml_server_t* s, int enable);
/* Queues a deck (the reference's JSON input) to run for its ceil(tLim/dt) steps. */
int chiml_server_submit(chiml_server_t* s, const char* deck_json, long long* id);
/* State of a request and its times in seconds since submission: admitted
 * (build start) and finished (last step completed on the device); -1 if not yet. */
int chiml_server_poll(chiml_server_t* s, long long id, int* state, double* admitted_s, double* finished_s);
/* Hands a finished simulation to the caller, who reads it through chiml_sim_*
 * and destroys it; its memory leaves the server's budget at this call. */
int chiml_server_take(chiml_server_t* s, long long id, chiml_sim_t** sim);
/* Failure message of a request (state CHIML_REQ_FAILED). */
int chiml_server_error(chiml_server_t* s, long long id, char* buf, size_t cap);
/* Blocks until nothing is queued, building or running.  Finished results
 * count against the budget until they are taken; when nothing is running the
 * next queued request is admitted regardless, so drain() completes even when
 * untaken results fill the budget (the device then holds more than it). */
int chiml_server_drain(chiml_server_t* s);
/* Totals: cell-updates completed, bytes held, requests active/queued now and
 * the largest number of simulations that were stepping at the same time. */
int chiml_server_stats(chiml_server_t* s, double* cell_updates, size_t* bytes_held, int* active, int* queued, int* peak_active);
void chiml_server_destroy(chiml_server_t* s);

#ifdef __cplusplus
}
#endif
#endif /* CHIML_SERVER_H */
