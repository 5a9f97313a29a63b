"""Hand-derived CIGARs of pairs whose optimal alignments tie.

Each case is built so that exactly one comparison of WFA2-lib's backtrace
priority decides the CIGAR.  The backtrace (smarco/WFA2-lib
wavefront_backtrace_affine; the reference's submodule, .gitmodules:1-3, is
empty, so its version is unpinned) starts at the end cell (diagonal
k = |t| - |p|, offset |t|) of the final score and repeats:

  1. in M, the candidates are the sources of the cell, each as an offset
       mismatch  M[s-x][k] + 1          I_open  M[s-o-e][k-1] + 1
       I_ext     I[s-e][k-1] + 1        D_open  M[s-o-e][k+1]
       D_ext     D[s-e][k+1]
     (in I only I_open / I_ext, in D only D_open / D_ext);
  2. the winner is the largest offset; equal offsets are decided by the
     type priority M 9 > D_ext 6 > D_open 5 > I_ext 2 > I_open 1, the
     backtrace_type values the reference restates in common.h:67-77 (the
     "piggy-backed" max: offset << 4 | type);
  3. in M, the matches between the cell's offset and the winner's are
     emitted first, then the winner's operation (X, I or D), and the walk
     moves to the winner's cell (score s-x, s-o-e or s-e; an *_ext winner
     continues in I or D, an *_open or mismatch winner in M);
  4. the walk ends at score 0, or in M when it reaches v = 0 or h = 0;
     then min(v, h) leading matches, v leading deletions and h leading
     insertions are emitted.

Notation below: v = pattern bases consumed, h = text bases consumed,
diagonal k = h - v, offset = h; "I" consumes text, "D" consumes pattern.
A wavefront value is the furthest offset on its diagonal reached at exactly
that score; each is given with the partial alignment that reaches it.
Every case names the tied candidates, the CIGAR the priority picks and the
CIGAR the other choice gives (co-optimal: same cost).  Away from the named
tie, every step of these walks has a single candidate at the largest
offset.

Prefix property (checked by the tests): for any sequence S, the pair
(S + p, S + t) has the CIGAR |S|M followed by the pair's own CIGAR, because
every wavefront cell of the longer pair is the shorter pair's cell shifted
by |S| along diagonal 0's extension through S.  The GPU tests use it to run
these cases through the kernels that the pair length selects.
"""

# (pattern, text, (x, o, e), expected CIGAR, tied candidates, CIGAR of the other choice)
TIES_462 = [
    # p=AA, t=C, score 12 = X 4 + D 8.  End cell k=-1, offset 1.
    #   mismatch M[8][-1]+1 = 0+1 = 1   (M[8][-1]: delete p[0] -> (1,0) at 8; p[1]=A != t[0]=C)
    #   D_open   M[4][0]    = 1         (M[4][0]: mismatch A/C -> (1,1) at 4)
    # Tie at 1, no match to emit.  M (9) > D_open (5): emit X, go to
    # M[8][-1], offset 0: (v,h) = (1,0), the walk ends, one leading D.
    ("AA", "C", (4, 6, 2), "1D1X", ("M", "D_open"), "1X1D"),
    # p=A, t=CC, score 12 = I 8 + X 4.  End cell k=1, offset 2.
    #   mismatch M[8][1]+1 = 1+1 = 2    (M[8][1]: insert t[0] -> (0,1) at 8; A != t[1]=C)
    #   I_open   M[4][0]+1 = 1+1 = 2    (M[4][0]: mismatch A/C -> (1,1))
    # Tie; M > I_open: X, then M[8][1] at (0,1): one leading I.
    ("A", "CC", (4, 6, 2), "1I1X", ("M", "I_open"), "1X1I"),
    # p=A, t=CAC, score 14 = I(2) 10 + X 4.  End cell k=2, offset 3.
    #   mismatch M[10][2]+1 = 2+1 = 3   (insert CA -> (0,2) at 10; A != t[2]=C)
    #   I_ext    I[12][1]+1 = 2+1 = 3   (I[12][1]: mismatch A/C -> (1,1) at 4, insert t[1] -> (1,2) at 12)
    # Tie; M (9) > I_ext (2): X, then (0,2) at 10: two leading I.
    ("A", "CAC", (4, 6, 2), "2I1X", ("M", "I_ext"), "1X2I"),
    # p=AAA, t=C, score 14 = D(2) 10 + X 4.  End cell k=-2, offset 1.
    #   mismatch M[10][-2]+1 = 0+1 = 1  (delete AA -> (2,0) at 10; A != C)
    #   D_ext    D[12][-1]   = 1        (mismatch A/C -> (1,1) at 4, delete p[1] -> (2,1) at 12)
    # Tie; M (9) > D_ext (6): X, then (2,0) at 10: two leading D.
    ("AAA", "C", (4, 6, 2), "2D1X", ("M", "D_ext"), "1X2D"),
    # p=ACG, t=CGAC, score 18 = I(2) 10 + D 8.  End cell k=1, offset 4.
    #   D_open   M[10][2]   = 4         (insert CG -> (0,2) at 10, match AC -> (2,4))
    #   I_ext    I[16][0]+1 = 3+1 = 4   (delete A -> (1,0) at 8, match CG -> (3,2),
    #                                    insert t[2] -> (3,3) at 16)
    # Tie at 4, no match.  D_open (5) > I_ext (2): D, then M[10][2] at
    # offset 4, whose only source is I_ext I[8][1]+1 = 2: 2 M down to
    # offset 2, where v = 0 ends the walk with two leading I.
    ("ACG", "CGAC", (4, 6, 2), "2I2M1D", ("D_open", "I_ext"), "1D2M2I"),
    # p=ACGA, t=GAC, score 18 = I 8 + D(2) 10.  End cell k=-1, offset 3.
    #   D_ext    D[16][0]    = 3        (insert G -> (0,1) at 8, match AC -> (2,3),
    #                                    delete p[2] -> (3,3) at 16)
    #   I_open   M[10][-2]+1 = 2+1 = 3  (delete AC -> (2,0) at 10, match GA -> (4,2))
    # Tie; D_ext (6) > I_open (1): D, continue in D[16][0], whose only
    # source is D_open M[8][1] = 3: D; then M[8][1] at offset 3: 2 M down to
    # offset 1, where v = 0 ends the walk with one leading I.
    ("ACGA", "GAC", (4, 6, 2), "1I2M2D", ("D_ext", "I_open"), "2D2M1I"),
    # p=CA, t=ACACAC, score 20 = I 8 + I(3) 12.  End cell k=4, offset 6.
    #   I_ext    I[18][3]+1 = 5+1 = 6   (insert A -> (0,1) at 8, match CA -> (2,3),
    #                                    insert AC -> (2,5) at 18)
    #   I_open   M[12][3]+1 = 5+1 = 6   (insert ACA -> (0,3) at 12, match CA -> (2,5))
    # Tie; I_ext (2) > I_open (1): the last gap is the long one.
    ("CA", "ACACAC", (4, 6, 2), "1I2M3I", ("I_ext", "I_open"), "3I2M1I"),
    # p=CACACA, t=AC, score 20 = D 8 + D(3) 12.  End cell k=-4, offset 2.
    #   D_ext    D[18][-3]  = 2         (delete C -> (1,0) at 8, match AC -> (3,2),
    #                                    delete AC -> (5,2) at 18)
    #   D_open   M[12][-3]  = 2         (delete CAC -> (3,0) at 12, match AC -> (5,2))
    # Tie; D_ext (6) > D_open (5).
    ("CACACA", "AC", (4, 6, 2), "1D2M3D", ("D_ext", "D_open"), "3D2M1D"),
    # p=CAC, t=GGACA, score 20.  End cell k=2, offset 5.
    #   D_open   M[12][3]   = 5         (insert GGA -> (0,3) at 12, match CA -> (2,5))
    #   I_open   M[12][1]+1 = 4+1 = 5   (insert G -> (0,1) at 8, mismatch C/G -> (1,2) at 12,
    #                                    match AC -> (3,4))
    #   (mismatch M[16][2]+1 = 3 and I_ext I[18][1]+1 = 3 are lower)
    # Tie at 5; D_open (5) > I_open (1).
    ("CAC", "GGACA", (4, 6, 2), "3I2M1D", ("D_open", "I_open"), "1I1X2M1I"),
    # p=AACG, t=CGGAA, score 22 = I(3) 12 + D(2) 10.  End cell k=1, offset 5.
    #   D_ext    D[20][2]   = 5         (insert CGG -> (0,3) at 12, match AA -> (2,5),
    #                                    delete C -> (3,5) at 20)
    #   I_ext    I[20][0]+1 = 4+1 = 5   (delete AA -> (2,0) at 10, match CG -> (4,2),
    #                                    insert GA -> (4,4) at 20)
    #   (D_open M[14][2] = 4 and mismatch M[18][1]+1 = 3 are lower)
    # Tie; D_ext (6) > I_ext (2).
    ("AACG", "CGGAA", (4, 6, 2), "3I2M2D", ("D_ext", "I_ext"), "2D2M3I"),
]

# The same ten comparisons under x=5, o=3, e=1 (a gap of L costs 3 + L),
# penalties that do not divide to 2/3/1 and so run on the LDS kernel.
TIES_531 = [
    # p=AC, t=GAA, score 9 = I 4 + X 5.  End cell k=1, offset 3.
    #   mismatch M[4][1]+1 = 2+1 = 3    (insert G -> (0,1) at 4, match A -> (1,2); C != A)
    #   D_open   M[5][2]   = 3          (insert GA -> (0,2) at 5, match A -> (1,3))
    # Tie; M > D_open: X, then M[4][1] at offset 2: 1 M down to offset 1,
    # where v = 0 ends the walk with one leading I.
    ("AC", "GAA", (5, 3, 1), "1I1M1X", ("M", "D_open"), "2I1M1D"),
    # p=AA, t=CAC, score 9.  End cell k=1, offset 3.
    #   mismatch M[4][1]+1 = 2+1 = 3    (insert C, match A -> (1,2); A != C)
    #   I_open   M[5][0]+1 = 2+1 = 3    (mismatch A/C -> (1,1) at 5, match A -> (2,2))
    # Tie; M > I_open.
    ("AA", "CAC", (5, 3, 1), "1I1M1X", ("M", "I_open"), "1X1M1I"),
    # p=ACG, t=CA, score 9.  End cell k=-1, offset 2.
    #   mismatch M[4][-1]+1 = 1+1 = 2   (delete A -> (1,0) at 4, match C -> (2,1); G != A)
    #   D_ext    D[8][0]    = 2         (insert C -> (0,1) at 4, match A -> (1,2),
    #                                    delete C -> (2,2) at 8)
    # Tie; M (9) > D_ext (6).
    ("ACG", "CA", (5, 3, 1), "1D1M1X", ("M", "D_ext"), "1I1M2D"),
    # p=AC, t=CAG, score 9.  End cell k=1, offset 3.
    #   mismatch M[4][1]+1 = 2+1 = 3    (insert C -> (0,1) at 4, match A -> (1,2); C != G)
    #   I_ext    I[8][0]+1  = 2+1 = 3   (delete A -> (1,0) at 4, match C -> (2,1),
    #                                    insert A -> (2,2) at 8)
    # Tie; M (9) > I_ext (2).
    ("AC", "CAG", (5, 3, 1), "1I1M1X", ("M", "I_ext"), "1D1M2I"),
    # p=AAA, t=CA, score 9.  End cell k=-1, offset 2.
    #   D_ext    D[8][0]    = 2         (insert C -> (0,1) at 4, match A -> (1,2),
    #                                    delete A -> (2,2) at 8)
    #   D_open   M[5][0]    = 2         (mismatch A/C -> (1,1) at 5, match A -> (2,2))
    #   (mismatch M[4][-1]+1 = 1 and I_open M[5][-2]+1 = 1 are lower)
    # Tie; D_ext (6) > D_open (5).
    ("AAA", "CA", (5, 3, 1), "1I1M2D", ("D_ext", "D_open"), "1X1M1D"),
    # p=AC, t=CGA, score 9.  End cell k=1, offset 3.
    #   D_open   M[5][2]    = 3         (insert CG -> (0,2) at 5, match A -> (1,3))
    #   I_ext    I[8][0]+1  = 2+1 = 3   (delete A -> (1,0) at 4, match C -> (2,1),
    #                                    insert G -> (2,2) at 8)
    # Tie; D_open (5) > I_ext (2).
    ("AC", "CGA", (5, 3, 1), "2I1M1D", ("D_open", "I_ext"), "1D1M2I"),
    # p=ACG, t=GA, score 9.  End cell k=-1, offset 2.
    #   D_ext    D[8][0]     = 2        (insert G -> (0,1) at 4, match A -> (1,2),
    #                                    delete C -> (2,2) at 8)
    #   I_open   M[5][-2]+1  = 1+1 = 2  (delete AC -> (2,0) at 5, match G -> (3,1))
    # Tie; D_ext (6) > I_open (1).
    ("ACG", "GA", (5, 3, 1), "1I1M2D", ("D_ext", "I_open"), "2D1M1I"),
    # p=A, t=CAAC, score 9.  End cell k=3, offset 4.
    #   I_ext    I[8][2]+1  = 3+1 = 4   (insert C -> (0,1) at 4, match A -> (1,2),
    #                                    insert A -> (1,3) at 8)
    #   I_open   M[5][2]+1  = 3+1 = 4   (insert CA -> (0,2) at 5, match A -> (1,3))
    # Tie; I_ext (2) > I_open (1).
    ("A", "CAAC", (5, 3, 1), "1I1M2I", ("I_ext", "I_open"), "2I1M1I"),
    # p=AC, t=CA, score 8.  End cell k=0, offset 2.
    #   D_open   M[4][1]    = 2         (insert C -> (0,1) at 4, match A -> (1,2))
    #   I_open   M[4][-1]+1 = 1+1 = 2   (delete A -> (1,0) at 4, match C -> (2,1))
    #   (no mismatch source: M[3] is empty, 3 < o+e and 3 < x)
    # Tie; D_open (5) > I_open (1).
    ("AC", "CA", (5, 3, 1), "1I1M1D", ("D_open", "I_open"), "1D1M1I"),
    # p=AA, t=CCC, score 11 = I(3) 6 + D(2) 5.  End cell k=1, offset 3.
    #   D_ext    D[10][2]   = 3         (insert CCC -> (0,3) at 6, delete A -> (1,3) at 10)
    #   I_ext    I[10][0]+1 = 2+1 = 3   (delete AA -> (2,0) at 5, insert CC -> (2,2) at 10)
    # Tie; D_ext (6) > I_ext (2).
    ("AA", "CCC", (5, 3, 1), "3I2D", ("D_ext", "I_ext"), "2D3I"),
]

# Gap placement in homopolymer runs (4/6/2): no priority is involved (one
# candidate), the rule is "largest offset first", which emits the matches
# nearest the end first and so puts the gap at the right end of the run.
# (pattern, text, (x, o, e), expected CIGAR, where, CIGAR of leftmost placement)
HOMOPOLYMER_462 = [
    # run at the start: M[0][0] = 3 (AAA match, then A != C).  End k=-1,
    # offset 4 at score 8: D_open M[0][0] = 3 is the only source: 1 M
    # (the C), D, then M[0][0] at offset 3: 3 leading M.
    ("AAAAC", "AAAC", (4, 6, 2), "3M1D1M", "start", "1D4M"),
    # run in the middle: M[0][0] = 4 (CAAA), D_open to k=-1 at 8 extends
    # over G: 1 M, D, 4 leading M.
    ("CAAAAG", "CAAAG", (4, 6, 2), "4M1D1M", "middle", "1M1D4M"),
    # run at the end: M[0][0] = 4 reaches the end of t; D_open at 8 is the
    # end cell itself (offset 4): no match, D, 4 leading M.
    ("CAAAA", "CAAA", (4, 6, 2), "4M1D", "end", "1M1D3M"),
    # an insertion in a run: M[0][0] = 4, I_open M[0][0]+1 = 5 at k=1,
    # extended over G to 6: 1 M, I, 4 leading M.
    ("CAAAG", "CAAAAG", (4, 6, 2), "4M1I1M", "middle", "1M1I4M"),
    # a gap before the run: M[0][0] = 0 (G != A); end k=-1 at score 8 from
    # D_open M[0][0] = 0 extended to 3: 3 M, then (v,h) = (1,0) ends the walk
    # with one leading D.
    ("GAAA", "AAA", (4, 6, 2), "1D3M", "start of the pair", "1D3M"),
]

ALL_TIES = TIES_462 + TIES_531


# Gap-affine 2-piece (a gap of length L costs gap(L) = min(o1 + e1 L,
# o2 + e2 L)).  The walk is the same with the candidates of both pieces:
# in M the mismatch, I1/I2 open M[s-o-e][k-1] + 1 and extend I[s-e][k-1] + 1,
# D1/D2 open M[s-o-e][k+1] and extend D[s-e][k+1] (o, e of the piece), in
# an I1 / D1 / I2 / D2 chain only that component's open and extend; equal
# offsets go to the larger type of common.h:67-77: M 9 > D2_ext 8 >
# D2_open 7 > D1_ext 6 > D1_open 5 > I2_ext 4 > I2_open 3 > I1_ext 2 >
# I1_open 1.  One case per adjacent pair of that order, so every other
# order inverts the comparison of at least one case.  Scores are the
# unscaled penalties of each case.
#
# (pattern, text, (x, o1, e1, o2, e2), expected CIGAR, tied candidates, CIGAR of the other choice)
TIES_2P = [
    # gap(1..4) = 4, 7, 10, 11 (L=3: piece 1 and 2 both 10).  p=CCGG, t=A,
    # score 14 = D(3) 10 + X 4.  End cell k=-3, offset 1.
    #   mismatch M[10][-3]+1 = 0+1 = 1  (delete CCG -> (3,0) at 10; G != A)
    #   D2_ext   D2[13][-2]  = 1        (mismatch C/A -> (1,1) at 4, D2 open 12 -> (2,1), 13 -> (3,1))
    #   D1_ext   D1[11][-2]  = 1        (from (1,1) at 4: D1 open 8 -> (2,1), 11 -> (3,1))
    # Tie at 1; M (9) > D2_ext (8): X, then (3,0) at 10: three leading D.
    # D2_ext first would follow the D2 chain back to the mismatch at (1,1).
    ("CCGG", "A", (4, 1, 3, 7, 1), "3D1X", ("M", "D2_ext"), "1X3D"),
    # gap(1..4) = 5, 8, 11, 12 (L=3 ties).  p=AAAAA, t=CA, score 17 = I 5 +
    # D(4) 12 = X 6 + D(3) 11.  End cell k=-3, offset 2: D2_ext D2[16][-2] = 2
    # (the D2 chains after 1I1M, (1,2) at 5, and after 1X1M, (2,2) at 6, both
    # reach (4,2) at 16) beats D1_ext D1[14][-2] = 2 (8 > 6), no match.  In
    # D2[16][-2] only D2_ext D2[15][-1] = 2 exists.  In D2[15][-1]:
    #   D2_ext   D2[14][0] = 2           (1I1M -> (1,2) at 5, D2 open 14 -> (2,2))
    #   D2_open  M[6][0]   = 2           (1X1M -> (2,2) at 6)
    # Tie; D2_ext (8) > D2_open (7): D2[14][0], whose only source is D2_open
    # M[5][1] = 2 ((0,1) at 5 extended A/A); in M[5][1] I1_open M[0][0]+1 = 1:
    # one M, then v = 0: one leading I.
    ("AAAAA", "CA", (6, 2, 3, 8, 1), "1I1M4D", ("D2_ext", "D2_open"), "1X1M3D"),
    # gap(1..7) = 9, 13, 17, 21, 25, 29, 33 (L=1: both pieces 9).  p=CAACAAC,
    # t=A: t's A against p[1], p[2], p[4] or p[5], each 34 (9+25, 13+21,
    # 21+13, 25+9).  End cell k=-6, offset 1:
    #   D2_open  M[25][-5] = 1           (delete CAACA -> (5,0) at 25, extend A/A -> (6,1))
    #   D1_ext   D1[30][-5] = 1          (the piece-1 chains after p[1], p[2] or p[4] reach (6,1) at 30)
    #   D1_open  M[25][-5] = 1
    # Tie; D2_open (7) > D1_ext (6) > D1_open (5): D, then M[25][-5] at
    # offset 1, pre-extension 0: one M, then (5,0): five leading D.
    # D1_ext first would stay in the piece-1 chain (D1_ext 6 > D1_open 5 at
    # each of its cells) back to p[1].
    ("CAACAAC", "A", (8, 5, 4, 4, 5), "5D1M1D", ("D2_open", "D1_ext"), "1D1M5D"),
    # gap(1..4) = 4, 6, 8, 10.  p=ACCG, t=C, score 10 = D 4 + D(2) 6 either way.
    # End cell k=-3, offset 1:
    #   D1_ext   D1[8][-2] = 1           (delete A -> (1,0) at 4, match C -> (2,1), D1 open 8 -> (3,1))
    #   D1_open  M[6][-2]  = 1           (delete AC -> (2,0) at 6, match C -> (3,1))
    # Tie; D1_ext (6) > D1_open (5): D1[8][-2], whose only source is D1_open
    # M[4][-1] = 1; there one M down to (1,0): one leading D.
    ("ACCG", "C", (4, 2, 2, 6, 1), "1D1M2D", ("D1_ext", "D1_open"), "2D1M1D"),
    # gap(1..4) = 4, 7, 10, 11 (L=3 ties).  p=AG, t=GCCA, score 14 = I(3) 10 +
    # D 4 = D 4 + I(3) 10.  End cell k=2, offset 4:
    #   D1_open  M[10][3]   = 4          (insert GCC -> (0,3) at 10, match A -> (1,4))
    #   I2_ext   I2[13][1]+1 = 3+1 = 4   (delete A -> (1,0) at 4, match G -> (2,1),
    #                                     I2 open 12 -> (2,2), 13 -> (2,3))
    #   I1_ext   I1[11][1]+1 = 3+1 = 4   (from (2,1) at 4: I1 open 8 -> (2,2), 11 -> (2,3))
    # Tie; D1_open (5) > I2_ext (4): D, then M[10][3] at offset 4, whose
    # sources I2_ext I2[9][2]+1 and I1_ext I1[7][2]+1 both give 3 (I2 first):
    # one M, then the insertion chain back to (0,0).
    ("AG", "GCCA", (4, 1, 3, 7, 1), "3I1M1D", ("D1_open", "I2_ext"), "1D1M3I"),
    # The insertion mirror of the D2 case above: gap(1..4) = 5, 8, 11, 12.
    # p=AG, t=GGCCC, score 17 = D 5 + I(4) 12 = X 6 + I(3) 11.  End cell k=3,
    # offset 5: I2_ext I2[16][2]+1 = 5 beats I1_ext I1[14][2]+1 = 5 (4 > 2);
    # in I2[16][2] only I2_ext I2[15][1]+1 = 4.  In I2[15][1]:
    #   I2_ext   I2[14][0]+1 = 2+1 = 3   (1D1M -> (2,1) at 5, I2 open 14 -> (2,2))
    #   I2_open  M[6][0]+1   = 2+1 = 3   (1X1M -> (2,2) at 6)
    # Tie; I2_ext (4) > I2_open (3): I2[14][0], from I2_open M[5][-1]+1 = 2;
    # in M[5][-1] one M down to (1,0): one leading D.
    ("AG", "GGCCC", (6, 2, 3, 8, 1), "1D1M4I", ("I2_ext", "I2_open"), "1X1M3I"),
    # gap(1..5) = 7, 11, 15, 19, 23 (L=1: both pieces 7).  p=C, t=ACACA,
    # score 22 = I 7 + I(3) 15 either way.  End cell k=4, offset 5:
    #   I2_open  M[15][3]+1  = 4+1 = 5   (insert ACA -> (0,3) at 15, match C -> (1,4))
    #   I1_ext   I1[18][3]+1 = 4+1 = 5   (insert A -> (0,1) at 7, match C -> (1,2),
    #                                     I1 open 14 -> (1,3), 18 -> (1,4))
    #   I1_open  M[15][3]+1  = 5
    # Tie; I2_open (3) > I1_ext (2): I, then M[15][3] at offset 4,
    # pre-extension 3: one M, then (0,3): three leading I.
    ("C", "ACACA", (4, 3, 4, 2, 5), "3I1M1I", ("I2_open", "I1_ext"), "1I1M3I"),
    # The insertion mirror of the D1 case: gap(1..4) = 4, 6, 8, 10.  p=A,
    # t=CAAG, score 10.  End cell k=3, offset 4:
    #   I1_ext   I1[8][2]+1 = 3+1 = 4    (insert C -> (0,1) at 4, match A -> (1,2), I1 open 8 -> (1,3))
    #   I1_open  M[6][2]+1  = 3+1 = 4    (insert CA -> (0,2) at 6, match A -> (1,3))
    # Tie; I1_ext (2) > I1_open (1): I1[8][2], from I1_open M[4][1]+1 = 3;
    # in M[4][1] one M down to (0,1): one leading I.
    ("A", "CAAG", (4, 2, 2, 6, 1), "1I1M2I", ("I1_ext", "I1_open"), "2I1M1I"),
]
