1 process(es),   1 device(s)
Bind solver rank 0 to device 0.
N_cell: 11630
N_dir: 32
N_band: 1
num_theta: 8
num_phi: 8

Iteration #0	 Margin per band per cell: 1.17287e+11
Time taken by inner loop: 696.245 milliseconds
Iteration #1	 Margin per band per cell: 6.00575e+10
Time taken by inner loop: 800.894 milliseconds
Iteration #2	 Margin per band per cell: 1.15765e+10
Time taken by inner loop: 782.221 milliseconds
Iteration #3	 Margin per band per cell: 4.01365e+09
Time taken by inner loop: 774.583 milliseconds
Iteration #4	 Margin per band per cell: 8.97455e+08
Time taken by inner loop: 770.681 milliseconds
Iteration #5	 Margin per band per cell: 5.11012e+08
Time taken by inner loop: 766.215 milliseconds
Iteration #6	 Margin per band per cell: 3.84585e+08
Time taken by inner loop: 764.253 milliseconds
Iteration #7	 Margin per band per cell: 2.87195e+08
Time taken by inner loop: 764.771 milliseconds
Iteration #8	 Margin per band per cell: 2.08553e+08
Time taken by inner loop: 765.422 milliseconds
Iteration #9	 Margin per band per cell: 1.33969e+08
Time taken by inner loop: 761.065 milliseconds
Iteration #10	 Margin per band per cell: 6.66569e+07
Time taken by inner loop: 763.65 milliseconds
Iteration #11	 Margin per band per cell: 2.41703e+07
Time taken by inner loop: 764.494 milliseconds
Iteration #12	 Margin per band per cell: 6.56251e+06
Time taken by inner loop: 768.542 milliseconds
Iteration #13	 Margin per band per cell: 1.46186e+06
Time taken by inner loop: 764.45 milliseconds
Iteration #14	 Margin per band per cell: 315434
Time taken by inner loop: 764.377 milliseconds
Iteration #15	 Margin per band per cell: 83936
Time taken by inner loop: 762.209 milliseconds
Iteration #16	 Margin per band per cell: 31599.8
Time taken by inner loop: 762.982 milliseconds
Iteration #17	 Margin per band per cell: 15060.6
Time taken by inner loop: 762.255 milliseconds
Iteration #18	 Margin per band per cell: 7725.16
Time taken by inner loop: 764.914 milliseconds
Iteration #19	 Margin per band per cell: 3876.38
Time taken by inner loop: 761.493 milliseconds
Iteration #20	 Margin per band per cell: 1811.56
Time taken by inner loop: 762.61 milliseconds
Iteration #21	 Margin per band per cell: 769.676
Time taken by inner loop: 763.79 milliseconds
Iteration #22	 Margin per band per cell: 295.636
Time taken by inner loop: 779.905 milliseconds
Iteration #23	 Margin per band per cell: 103.904
Time taken by inner loop: 812.668 milliseconds
Iteration #24	 Margin per band per cell: 34.4504
Time taken by inner loop: 762.143 milliseconds
Iteration #25	 Margin per band per cell: 11.301
Time taken by inner loop: 764.076 milliseconds
Iteration #26	 Margin per band per cell: 3.86679
Time taken by inner loop: 764.343 milliseconds
Iteration #27	 Margin per band per cell: 1.42699
Time taken by inner loop: 766.21 milliseconds
Iteration #28	 Margin per band per cell: 0.566166
Time taken by inner loop: 762.407 milliseconds
Iteration #29	 Margin per band per cell: 0.234507
Time taken by inner loop: 766.925 milliseconds
Iteration #30	 Margin per band per cell: 0.0981061
Time taken by inner loop: 763.02 milliseconds
Iteration #31	 Margin per band per cell: 0.0404515
Time taken by inner loop: 762.267 milliseconds
Iteration #32	 Margin per band per cell: 0.0162195
Time taken by inner loop: 762.11 milliseconds
Iteration #33	 Margin per band per cell: 0.00629932
Time taken by inner loop: 763.066 milliseconds
Iteration #34	 Margin per band per cell: 0.00237878
Time taken by inner loop: 766.555 milliseconds
Iteration #35	 Margin per band per cell: 0.000882921
Time taken by inner loop: 765.026 milliseconds
Iteration #36	 Margin per band per cell: 0.000326491
Time taken by inner loop: 760.728 milliseconds
Iteration #37	 Margin per band per cell: 0.000122008
Time taken by inner loop: 762.741 milliseconds
Iteration #38	 Margin per band per cell: 4.63701e-05
Time taken by inner loop: 762.648 milliseconds


Time taken by iteration: 29861.1 milliseconds
heat flux of Boundary Condition #0: -1.29677e+09
heat flux of Boundary Condition #1: 873.006
heat flux of Boundary Condition #2: -1264.5
heat flux of Boundary Condition #3: 1.29677e+09
Time taken by solver: 39637.6 milliseconds
