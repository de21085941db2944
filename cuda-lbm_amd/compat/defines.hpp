// defines.hpp — compile-time configuration seen by scenario headers (src/defines.hpp:1-72).
// This build is D2Q9 only. NX, NY and SCALE may be given on the command line; otherwise the
// reference's per-scenario grid table applies (defines.hpp:20-35).
#ifndef DEFINES_H
#define DEFINES_H

#ifndef D2Q9
#define D2Q9
#endif

#if !defined(NX) || !defined(NY)
#if defined(USE_TAYLOR_GREEN)
#define SCALE 2
#define NX (128 * SCALE)
#define NY (128 * SCALE)
#elif defined(USE_POISEUILLE)
#define SCALE 1
#define NX (150 * SCALE)
#define NY (100 * SCALE)
#elif defined(USE_LID_DRIVEN)
#define SCALE 1
#define NX (129 * SCALE)
#define NY (129 * SCALE)
#elif defined(USE_FLOW_PAST_CYLINDER)
#define SCALE 1
#define NX (256 * SCALE)
#define NY (128 * SCALE)
#else
#error "define NX and NY (and optionally SCALE), or one of the USE_* scenario switches"
#endif
#endif

#ifndef SCALE
#define SCALE 1
#endif
#ifndef NZ
#define NZ 1
#endif
#define BLOCK_SIZE 16

#endif  // DEFINES_H
