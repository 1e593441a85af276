/* Public types of the mjmlx C API, HIP engine.
 *
 * Byte-identical to the reference's caller-allocated types so existing
 * callers link unchanged (reference include/mjmlx/mjmlx_types.h:27-89):
 * the visibility macro, the opaque handles, the integrator / solver /
 * disable-flag enums (MuJoCo mjtIntegrator / mjtSolver / mjtDisableBit
 * values), MjmlxModelInfo (9 ints) and MjmlxBatchedConfig (5 ints).
 */
#ifndef MJMLX_TYPES_H
#define MJMLX_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MJMLX_API __attribute__((visibility("default")))

/* opaque handles (reference mjmlx_types.h:34-37) */
typedef struct MjmlxModel MjmlxModel;
typedef struct MjmlxData MjmlxData;
typedef struct MjmlxBatchedSim MjmlxBatchedSim;
typedef struct MjmlxCameraRenderer MjmlxCameraRenderer;

/* reference mjmlx_types.h:40-45 */
typedef enum {
  MJMLX_INTEGRATOR_EULER = 0,
  MJMLX_INTEGRATOR_RK4 = 1,
  MJMLX_INTEGRATOR_IMPLICIT = 2,
  MJMLX_INTEGRATOR_IMPLICITFAST = 3,
} MjmlxIntegrator;

/* reference mjmlx_types.h:48-52 */
typedef enum {
  MJMLX_SOLVER_PGS = 0,
  MJMLX_SOLVER_CG = 1,
  MJMLX_SOLVER_NEWTON = 2,
} MjmlxSolver;

/* reference mjmlx_types.h:55-67 */
typedef enum {
  MJMLX_DISABLE_CONSTRAINT = 1 << 0,
  MJMLX_DISABLE_EQUALITY = 1 << 1,
  MJMLX_DISABLE_FRICTIONLOSS = 1 << 2,
  MJMLX_DISABLE_LIMIT = 1 << 3,
  MJMLX_DISABLE_CONTACT = 1 << 4,
  MJMLX_DISABLE_PASSIVE = 1 << 5,
  MJMLX_DISABLE_GRAVITY = 1 << 6,
  MJMLX_DISABLE_CLAMPCTRL = 1 << 7,
  MJMLX_DISABLE_WARMSTART = 1 << 8,
  MJMLX_DISABLE_FILTERPARENT = 1 << 9,
  MJMLX_DISABLE_ACTUATION = 1 << 10,
} MjmlxDisableFlag;

/* model dimensions (reference mjmlx_types.h:70-80) */
typedef struct {
  int nq;
  int nv;
  int nu;
  int na;
  int nbody;
  int njnt;
  int ngeom;
  int nsite;
  int ncon;
} MjmlxModelInfo;

/* batched config (reference mjmlx_types.h:83-89); use_gpu = 1 selects the
 * HIP engine; 0 asks for the MuJoCo C CPU path, which this engine does not
 * contain (create then fails with a message on stderr). */
typedef struct {
  int num_envs;
  MjmlxIntegrator integrator;
  int use_gpu;
  int solver_iterations;
  int domain_rand;
} MjmlxBatchedConfig;

#ifdef __cplusplus
}
#endif

#endif /* MJMLX_TYPES_H */
