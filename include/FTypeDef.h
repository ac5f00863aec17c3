/* Floating-point type of the ParODE C ABI (Include/FTypeDef.h:1-7 of the reference). */
#ifndef __FTYPE__
#define __FTYPE__

typedef float fType;

#endif
