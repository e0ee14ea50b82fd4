// gl25.h -- written by gen_gl25.py: 25-point Gauss-Legendre rule on [-1, 1],
// nodes ascending, each value the double nearest to the exact one.
#pragma once
// Initializer lists, for host and device arrays alike.
#define AMC_GL25_N 25
#define AMC_GL25_X { \
    -0x1.fdb9a4a51c2f7p-1, \
    -0x1.f40d4b23b4175p-1, \
    -0x1.e2cd9020c6b0cp-1, \
    -0x1.ca3c642223461p-1, \
    -0x1.aab8fe033d775p-1, \
    -0x1.84bda14fdc6c2p-1, \
    -0x1.58ddb0e88edd5p-1, \
    -0x1.27c36f82fec31p-1, \
    -0x1.e45ad3f9e97d3p-2, \
    -0x1.71d72726b05d8p-2, \
    -0x1.f3707b1145cbdp-3, \
    -0x1.f740f7c98999ep-4, \
    0x0.0p+0, \
    0x1.f740f7c98999ep-4, \
    0x1.f3707b1145cbdp-3, \
    0x1.71d72726b05d8p-2, \
    0x1.e45ad3f9e97d3p-2, \
    0x1.27c36f82fec31p-1, \
    0x1.58ddb0e88edd5p-1, \
    0x1.84bda14fdc6c2p-1, \
    0x1.aab8fe033d775p-1, \
    0x1.ca3c642223461p-1, \
    0x1.e2cd9020c6b0cp-1, \
    0x1.f40d4b23b4175p-1, \
    0x1.fdb9a4a51c2f7p-1}
#define AMC_GL25_W { \
    0x1.755a1bf835241p-7, \
    0x1.afccd36646f49p-6, \
    0x1.4f5fa2650dc2fp-5, \
    0x1.c1c77e339dfb8p-5, \
    0x1.16af5d2993100p-4, \
    0x1.48419d7047213p-4, \
    0x1.74da0d03a8230p-4, \
    0x1.9bcb95550b7c7p-4, \
    0x1.bc7f12e17567fp-4, \
    0x1.d6759d2b4d1adp-4, \
    0x1.e94a725b474a6p-4, \
    0x1.f4b47ded39736p-4, \
    0x1.f8877426815bep-4, \
    0x1.f4b47ded39736p-4, \
    0x1.e94a725b474a6p-4, \
    0x1.d6759d2b4d1adp-4, \
    0x1.bc7f12e17567fp-4, \
    0x1.9bcb95550b7c7p-4, \
    0x1.74da0d03a8230p-4, \
    0x1.48419d7047213p-4, \
    0x1.16af5d2993100p-4, \
    0x1.c1c77e339dfb8p-5, \
    0x1.4f5fa2650dc2fp-5, \
    0x1.afccd36646f49p-6, \
    0x1.755a1bf835241p-7}
