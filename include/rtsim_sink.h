/*
 * rtsim_sink.h -- ping files of the rtsim C ABI.
 *
 * FileSink (sinks/FileSink.h, FileSink.cpp:8-42) appends every ping handed to
 * set_output() to one binary file through rtac_base's serialize(), whose
 * format is not part of the reference.  This library defines its own
 * stream of ping records (little endian):
 *
 *   offset  size  field
 *        0     4  magic "RTPG"
 *        4     4  u32 version = 1
 *        8     4  u32 kind: 1 = 1D (width 1), 2 = 2D
 *       12     4  u32 scalar: 1 = complex64
 *       16     4  u32 height (range bins)
 *       20     4  u32 width (bearings)
 *       24     4  f32 range_min (m)
 *       28     4  f32 range_max (m)
 *       32     8  u64 ping index in the file
 *       40     8  u64 FNV-1a 64 of the bearing table and the data
 *       48   4*W  f32 bearings (radians; 2D only, W = 0 for 1D)
 *        .  8*H*W complex64 data, row-major [range][bearing]
 */
#ifndef RTSIM_SINK_H
#define RTSIM_SINK_H

#include <stddef.h>
#include <stdint.h>

#include "rtsim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtsim_sink_s* rtsim_sink;

typedef struct {
    uint32_t kind, scalar, height, width;
    float range_min, range_max;
    uint64_t index;
    uint64_t checksum;
} rtsim_ping_header;

/* FileSink::Create(filename, overwrite) (FileSink.cpp:15-24): creates missing
 * parent directories; an existing file without overwrite is RTSIM_ERR_IO.     */
int rtsim_sink_open(const char* path, int overwrite, rtsim_sink* out);
/* FileSink::set_output(sensor) (FileSink.cpp:26-42): downloads the sensor's
 * ping (synchronising `stream`) and appends one record.                       */
int rtsim_sink_write(rtsim_sink sink, rtsim_sensor sensor, void* stream);
/* Appends one record from host data (bearings may be NULL for kind 1).        */
int rtsim_sink_write_host(rtsim_sink sink, int kind, uint32_t height, uint32_t width,
                          float range_min, float range_max, const float* bearings,
                          const float* data);
int rtsim_sink_count(rtsim_sink sink, uint64_t* pings);
int rtsim_sink_close(rtsim_sink sink);

/* Readers: number of records, and record `index` (bearings / data may be NULL
 * to read the header only; sizes from the header).  A record whose checksum
 * does not match is RTSIM_ERR_IO.                                             */
int rtsim_ping_file_count(const char* path, uint64_t* pings);
int rtsim_ping_file_read(const char* path, uint64_t index, rtsim_ping_header* header,
                         float* bearings, float* data);

#ifdef __cplusplus
}
#endif

#endif /* RTSIM_SINK_H */
