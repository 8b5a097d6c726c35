/* gmi_flops.h -- algorithmic flop counts per pair (frozen).
 *
 * Counted from the reference's far-field branch as written
 * (add/sub/mul/div/sqrt = 1 flop each, constants folded, branch selection
 * free).  Self pairs count, as the reference evaluates them in-loop.  These
 * are the numerators of every achieved-FLOP/s figure this project reports;
 * the kernels' own instruction counts do not enter.
 */
#ifndef GMI_FLOPS_H
#define GMI_FLOPS_H

#define GMI_FLOPS_SS 92 /* cuda_functions.cu:302-398, 6x6 seg<-seg       */
#define GMI_FLOPS_SB 56 /* cuda_functions.cu:593-667, seg (v,w) <- blob f */
#define GMI_FLOPS_BS 56 /* cuda_functions.cu:755-829, blob v <- seg f,tau */
#define GMI_FLOPS_BB 40 /* cuda_functions.cu:478-517, 3x3 blob<-blob     */
/* Above a no-slip wall, translations only (the PRESCRIBED_CILIA rows):
 * unbounded far form 40 + image corrections 97, common subexpressions
 * counted once */
#define GMI_FLOPS_SS_WALL_TT 137 /* cuda_functions.cu:114-177 */
/* Above a no-slip wall, 6x6: unbounded 92 (as ss) + image tt 97 + rr 38 +
 * rt 48 + tr 49 */
#define GMI_FLOPS_SS_WALL 324    /* cuda_functions.cu:114-213 */

#endif
