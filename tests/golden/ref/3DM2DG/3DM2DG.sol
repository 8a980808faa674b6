N_cell: 68
N_dir: 64
N_band: 1
num_theta: 4
num_phi: 4

Iteration #0	 Margin per band per cell: 1.96147e+11
Iteration #1	 Margin per band per cell: 7.77877e+10
Iteration #2	 Margin per band per cell: 1.06402e+10
Iteration #3	 Margin per band per cell: 3.45176e+09
Iteration #4	 Margin per band per cell: 1.72112e+09
Iteration #5	 Margin per band per cell: 7.44677e+08
Iteration #6	 Margin per band per cell: 2.856e+08
Iteration #7	 Margin per band per cell: 1.01549e+08
Iteration #8	 Margin per band per cell: 3.46919e+07
Iteration #9	 Margin per band per cell: 1.16711e+07
Iteration #10	 Margin per band per cell: 3.92402e+06
Iteration #11	 Margin per band per cell: 1.32793e+06
Iteration #12	 Margin per band per cell: 453155
Iteration #13	 Margin per band per cell: 155792
Iteration #14	 Margin per band per cell: 53851.9
Iteration #15	 Margin per band per cell: 18679.5
Iteration #16	 Margin per band per cell: 6491.91
Iteration #17	 Margin per band per cell: 2258.36
Iteration #18	 Margin per band per cell: 785.917
Iteration #19	 Margin per band per cell: 273.525
Iteration #20	 Margin per band per cell: 95.1919
Iteration #21	 Margin per band per cell: 33.1265
Iteration #22	 Margin per band per cell: 11.5272
Iteration #23	 Margin per band per cell: 4.01105
Iteration #24	 Margin per band per cell: 1.39566
Iteration #25	 Margin per band per cell: 0.485619
Iteration #26	 Margin per band per cell: 0.16897
Iteration #27	 Margin per band per cell: 0.0587922
Iteration #28	 Margin per band per cell: 0.0204564
Iteration #29	 Margin per band per cell: 0.00711763
Iteration #30	 Margin per band per cell: 0.00247651
Iteration #31	 Margin per band per cell: 0.00086167
Iteration #32	 Margin per band per cell: 0.000299805
Iteration #33	 Margin per band per cell: 0.000104312
Iteration #34	 Margin per band per cell: 3.62934e-05

heat flux of Boundary Condition #0: -9.92317e+08
heat flux of Boundary Condition #1: -113067
heat flux of Boundary Condition #2: -74043.1
heat flux of Boundary Condition #3: 9.92036e+08
