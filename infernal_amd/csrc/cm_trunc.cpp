/* cm_trunc.cpp -- truncated-begin penalties (host side).
 *
 * Restates cm_tr_penalties_Create (src/cm_trunc.c:51-450) and the expected
 * occupancy it depends on (cm_ExpectedStateOccupancy src/cm.c:4250,
 * cm_ExpectedPositionOccupancy :4405, InsertsGivenNodeIndex :4897).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "cm_internal.h"

extern "C" int infx_tr_penalties_create(CM_t *cm, char *errbuf)
{
  return infx_fail(eslEINVAL, errbuf, "truncation penalties are not available in this build");
}
