// Types and enums of the hpcscan host API.
//
// Names and enumerator values are the reference's API (src/type_def.h:1-112),
// kept identical so that a Grid backend written against this harness compiles
// against the reference tree and vice versa.
//
// Storage / compute precision is chosen at build time, as in the reference
// (src/type_def.h:21-43): -DHPCSCAN_STORAGE_FP64 selects double storage,
// -DHPCSCAN_COMPUTE_FP64 double compute.
#ifndef HPCSCAN_TYPE_DEF_H_
#define HPCSCAN_TYPE_DEF_H_

#include <string>

using namespace std;

namespace hpcscan {

typedef int      Myint;
typedef int      Myint32;
typedef long int Myint64;
typedef float    Myfloat32;
typedef double   Myfloat64;

#ifdef HPCSCAN_STORAGE_FP64
typedef Myfloat64 Myfloat;
#else
typedef Myfloat32 Myfloat;
#endif

#ifdef HPCSCAN_COMPUTE_FP64
typedef Myfloat64 Myfloat2;
#else
typedef Myfloat32 Myfloat2;
#endif

// Return codes used by all methods
enum Rtn_code {RTN_CODE_OK = 0, RTN_CODE_KO = -1, RTN_CODE_EXIT = -2};

enum Debug_level {NO_DEBUG = -1, LIGHT_DEBUG, MID_DEBUG, FULL_DEBUG};

enum Display_type {MASTER = 0, ALL = 1};

enum Axis_type {AXIS1, AXIS2, AXIS3};

enum Dim_type {DIM1 = 1, DIM2 = 2, DIM3 = 3};

enum FD_type {FD_D1, FD_D2};

enum Func_type {FUNC_NONE, FUNC_SINE, FUNC_COSINE, FUNC_CONST, FUNC_LINEAR, FUNC_DIRAC, FUNC_RANDOM};

enum Point_type {INNER_POINTS, ALL_POINTS, MIDDLE_POINT,
                 I1HALO1, I1HALO2, I2HALO1, I2HALO2, I3HALO1, I3HALO2,
                 I1INNERHALO1, I1INNERHALO2, I2INNERHALO1, I2INNERHALO2, I3INNERHALO1, I3INNERHALO2};

enum Grid_type {GRID_LOCAL, GRID_GLOBAL};

enum {UNSPECIFIED = -1};

enum MPI_comm_mode_type {MPI_COMM_MODE_SENDRECV};

enum BoundCond_type {NO_BOUND_COND, BOUND_COND_ANTI_MIRROR};

enum PropaInit_type {EIGEN_MODE, MODELING};

enum Modeling_type {FORWARD, ADJOINT};

enum DataAcqui_type {ACQUI_BUILTIN};

// hardware counter sample (src/type_def.h:46-55)
typedef struct {
    double timeOfMeasure;
    Myfloat powerWatt;
} hwCounter_struct_type;

} // namespace hpcscan

#endif
