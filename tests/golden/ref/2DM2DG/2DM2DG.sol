N_cell: 578
N_dir: 32
N_band: 1
num_theta: 8
num_phi: 8

Iteration #0	 Margin per band per cell: 1.12713e+11
Iteration #1	 Margin per band per cell: 5.14895e+10
Iteration #2	 Margin per band per cell: 8.32961e+09
Iteration #3	 Margin per band per cell: 2.40969e+09
Iteration #4	 Margin per band per cell: 7.01424e+08
Iteration #5	 Margin per band per cell: 3.21018e+08
Iteration #6	 Margin per band per cell: 2.26782e+08
Iteration #7	 Margin per band per cell: 1.59011e+08
Iteration #8	 Margin per band per cell: 1.03984e+08
Iteration #9	 Margin per band per cell: 6.28146e+07
Iteration #10	 Margin per band per cell: 3.52158e+07
Iteration #11	 Margin per band per cell: 1.8529e+07
Iteration #12	 Margin per band per cell: 9.26839e+06
Iteration #13	 Margin per band per cell: 4.46245e+06
Iteration #14	 Margin per band per cell: 2.09093e+06
Iteration #15	 Margin per band per cell: 962450
Iteration #16	 Margin per band per cell: 438646
Iteration #17	 Margin per band per cell: 199186
Iteration #18	 Margin per band per cell: 90558.7
Iteration #19	 Margin per band per cell: 41355.3
Iteration #20	 Margin per band per cell: 19007.7
Iteration #21	 Margin per band per cell: 8798.24
Iteration #22	 Margin per band per cell: 4100.67
Iteration #23	 Margin per band per cell: 1922.63
Iteration #24	 Margin per band per cell: 905.784
Iteration #25	 Margin per band per cell: 428.217
Iteration #26	 Margin per band per cell: 202.922
Iteration #27	 Margin per band per cell: 96.2892
Iteration #28	 Margin per band per cell: 45.7201
Iteration #29	 Margin per band per cell: 21.7111
Iteration #30	 Margin per band per cell: 10.308
Iteration #31	 Margin per band per cell: 4.89218
Iteration #32	 Margin per band per cell: 2.32087
Iteration #33	 Margin per band per cell: 1.10057
Iteration #34	 Margin per band per cell: 0.521726
Iteration #35	 Margin per band per cell: 0.247257
Iteration #36	 Margin per band per cell: 0.11716
Iteration #37	 Margin per band per cell: 0.0555096
Iteration #38	 Margin per band per cell: 0.0262992
Iteration #39	 Margin per band per cell: 0.0124601
Iteration #40	 Margin per band per cell: 0.00590373
Iteration #41	 Margin per band per cell: 0.00279743
Iteration #42	 Margin per band per cell: 0.00132564
Iteration #43	 Margin per band per cell: 0.000628243
Iteration #44	 Margin per band per cell: 0.000297757
Iteration #45	 Margin per band per cell: 0.000141131
Iteration #46	 Margin per band per cell: 6.68975e-05

heat flux of Boundary Condition #0: -1.29462e+09
heat flux of Boundary Condition #1: 645.383
heat flux of Boundary Condition #2: -30904.6
heat flux of Boundary Condition #3: 1.2946e+09
