/* Single-precision instance of the scalar-strategy restatement: the
   reference's `-D FLOAT32` build (butterfly/core/types.mojo:7).
   TEST INFRASTRUCTURE only, like bfly_oracle.c. */
#define FLOAT_T float
#define SFX(name) name##_f32
#include "bfly_oracle.c"
