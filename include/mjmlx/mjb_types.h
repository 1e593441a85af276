/* Public types of the mjb unified C API (HIP engine build).
 *
 * Byte-identical to the reference (reference include/mjmlx/mjb_types.h:27-65):
 * visibility macro, backend enum values, opaque handles, MjbModelInfo
 * (12 ints), MjbBatchedConfig (2 ints).
 */
#ifndef MJB_TYPES_H
#define MJB_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

#define MJB_API __attribute__((visibility("default")))

/* reference mjb_types.h:34-37: value 1 means "the GPU engine"; in this build
 * that engine is the HIP backend, so callers pass the same enum unchanged. */
typedef enum {
  MJB_BACKEND_CPU = 0,  /* MuJoCo C (double); needs libmujoco, absent from this engine */
  MJB_BACKEND_MLX = 1,  /* the GPU engine: HIP on gfx950 */
} MjbBackendType;
#define MJB_BACKEND_HIP MJB_BACKEND_MLX

typedef struct MjbBackend MjbBackend;
typedef struct MjbModel MjbModel;
typedef struct MjbData MjbData;
typedef struct MjbBatchedSim MjbBatchedSim;

/* reference mjb_types.h:46-59 */
typedef struct {
  int nq;
  int nv;
  int nu;
  int nbody;
  int njnt;
  int ngeom;
  int nsite;
  int nmocap;
  int ntendon;
  int nsensor;
  int nsensordata;
  int neq;
} MjbModelInfo;

/* reference mjb_types.h:62-65 */
typedef struct {
  int num_envs;
  int solver_iterations;
} MjbBatchedConfig;

#ifdef __cplusplus
}
#endif

#endif /* MJB_TYPES_H */
