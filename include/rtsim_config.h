/*
 * rtsim_config.h -- YAML sensor configuration of the rtsim C ABI.
 *
 * The drop-in for SensorFactory / SensorInfoFactory
 * (SensorInfoFactory.cpp:7-235) and FileFinder (factories/utilities.cpp:22-109):
 * the reference's receiver YAML schema is read as is (yaml-cpp is replaced by
 * a built-in reader of the subset the configs use, anchors/aliases included).
 */
#ifndef RTSIM_CONFIG_H
#define RTSIM_CONFIG_H

#include <stddef.h>

#include "rtsim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtsim_info_s* rtsim_info;

/* SensorInfoFactory::Make(path) (SensorInfoFactory.cpp:76-80): parses the
 * YAML into a sensor description (host only, no device work).  Errors:
 * RTSIM_ERR_IO (file not found), RTSIM_ERR_CONFIG (ConfigError cases:
 * missing keys, unknown type, bad units, ...).                                */
int rtsim_info_from_yaml(const char* path, rtsim_info* out);
/* The description; its pointers stay valid until rtsim_info_destroy.          */
int rtsim_info_desc(rtsim_info info, const rtsim_sensor_desc** desc);
int rtsim_info_destroy(rtsim_info info);

/* SensorFactory::Make(path) (SensorInfoFactory.cpp:54-57): description +
 * rtsim_sensor_create.                                                         */
int rtsim_sensor_create_from_yaml(const char* path, rtsim_sensor* out);

/* FileFinder::add_search_path / find_one (utilities.cpp:62-97).  The search
 * also covers RTAC_CONFIG_PATH (':'-separated) and the YAML file's folder.    */
int rtsim_config_add_search_path(const char* path);
int rtsim_config_find(const char* name, char* out, size_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* RTSIM_CONFIG_H */
