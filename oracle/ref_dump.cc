// ref_dump.cc -- full-precision output hook for the compiled reference.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile builds the unmodified
// LULESH/omp_4.0 sources (lulesh.cc, lulesh-init.cc, lulesh-comm.cc,
// lulesh-viz.cc, lulesh-util.cc) from /root/reference.  lulesh-util.cc is
// compiled with its VerifyAndWriteFinalOutput renamed to
// ref_VerifyAndWriteFinalOutput_text, and this file supplies the
// VerifyAndWriteFinalOutput that main (lulesh.cc:3267) calls: it writes the
// final state at full precision, then calls the reference's own printout.
//
// With LULESH_REF_DUMP=<path> set, the dump is a raw little-endian file:
//   int32 nx, int32 cycle, float64 time, float64 deltatime,
//   then per field (e p q v ss x y z xd yd zd) its float64 values.
// With LULESH_REF_REGIONS=<path> set, the region layout of
// CreateRegionIndexSets (lulesh-init.cc:377-484) and the elemRep of main
// (lulesh.cc:3171-3192) are written as int32: numElem, numReg, then
// regNumList[numElem] (1-based region of each element), elemRep[numElem],
// regElemSize[numReg].
#include <stdio.h>
#include <stdlib.h>
#include "lulesh.h"

void ref_VerifyAndWriteFinalOutput_text(Real_t elapsed_time, Domain& locDom,
                                        Int_t nx, Int_t numRanks);

static void put(FILE *f, const std::vector<Real_t> &v, size_t n)
{
   fwrite(v.data(), sizeof(Real_t), n, f);
}

void VerifyAndWriteFinalOutput(Real_t elapsed_time, Domain& locDom,
                               Int_t nx, Int_t numRanks)
{
   printf("FullPrecision OriginEnergy = %.17e (%a)\n", locDom.e(0), locDom.e(0));
   printf("FullPrecision Time = %.17e DeltaTime = %.17e Cycle = %d\n",
          (double)locDom.time(), (double)locDom.deltatime(), locDom.cycle());
   const char *path = getenv("LULESH_REF_DUMP");
   if (path != NULL && path[0] != '\0') {
      FILE *f = fopen(path, "wb");
      if (f != NULL) {
         int hdr[2] = { nx, locDom.cycle() };
         double t[2] = { locDom.time(), locDom.deltatime() };
         fwrite(hdr, sizeof(int), 2, f);
         fwrite(t, sizeof(double), 2, f);
         size_t ne = (size_t)locDom.numElem(), nn = (size_t)locDom.numNode();
         put(f, locDom.m_e, ne); put(f, locDom.m_p, ne); put(f, locDom.m_q, ne);
         put(f, locDom.m_v, ne); put(f, locDom.m_ss, ne);
         put(f, locDom.m_x, nn); put(f, locDom.m_y, nn); put(f, locDom.m_z, nn);
         put(f, locDom.m_xd, nn); put(f, locDom.m_yd, nn); put(f, locDom.m_zd, nn);
         fclose(f);
      }
   }
   const char *rpath = getenv("LULESH_REF_REGIONS");
   if (rpath != NULL && rpath[0] != '\0') {
      FILE *f = fopen(rpath, "wb");
      if (f != NULL) {
         int hdr[2] = { (int)locDom.numElem(), (int)locDom.numReg() };
         fwrite(hdr, sizeof(int), 2, f);
         fwrite(locDom.regNumList(), sizeof(Index_t), (size_t)locDom.numElem(), f);
         fwrite(locDom.m_elemRep.data(), sizeof(Index_t), (size_t)locDom.numElem(), f);
         for (Int_t r = 0; r < locDom.numReg(); ++r) {
            Index_t n = locDom.regElemSize(r);
            fwrite(&n, sizeof(Index_t), 1, f);
         }
         fclose(f);
      }
   }
   ref_VerifyAndWriteFinalOutput_text(elapsed_time, locDom, nx, numRanks);
}
