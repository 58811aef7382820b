// common_defines.h — pipeline constants of the reference host API
// (common/inc/common_defines.h), same names and values, for source
// compatibility of callers.
#ifndef KFX_DROPIN_COMMON_DEFINES
#define KFX_DROPIN_COMMON_DEFINES

#define DEFAULT_STRIDE 256
#define MAX_DEPTH_VALUE 32767
#define MAX_TSDF_WEIGHT (1 << 7)
#define DEFAULT_NUM_OF_ICP_PYRAMID_LEVELS 3
#define DEFAULT_SCALE_FACTOR (1.f / 1000.f)
#define DEFAULT_SIGMA_COLOR 30.f
#define DEFAULT_SIGMA_SPACE 4.5f
#define DEFAULT_MAX_DISTANCE 300
#define DEFAULT_TRUNCATION_DISTANCE 0.03f
#define DEFAULT_MAX_WEIGHT 1 /* colour volume */
#define DEFAULT_DISTANCE_THRESHOLD 0.1f
#define DEFAULT_CAMERA_MOVEMENT_THRESHOLD 0.001f
#define DEFAULT_VOLUME_SHIFTING_THRESHOLD 1.5f
#define DIVUP(TOTAL, DIVISOR) (((TOTAL) + (DIVISOR)-1) / DIVISOR)

#endif
