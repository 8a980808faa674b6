N_cell: 11630
N_dir: 32
N_band: 1
num_theta: 8
num_phi: 8

Iteration #0	 Margin per band per cell: 1.17287e+11
Time taken by inner loop: 38039 milliseconds
Iteration #1	 Margin per band per cell: 6.00573e+10
Time taken by inner loop: 38364 milliseconds
Iteration #2	 Margin per band per cell: 1.15764e+10
Time taken by inner loop: 37929.8 milliseconds
Iteration #3	 Margin per band per cell: 4.01362e+09
Time taken by inner loop: 38409.7 milliseconds
Iteration #4	 Margin per band per cell: 8.97319e+08
Time taken by inner loop: 38181 milliseconds
Iteration #5	 Margin per band per cell: 5.10944e+08
Time taken by inner loop: 38444.8 milliseconds
Iteration #6	 Margin per band per cell: 3.8453e+08
Time taken by inner loop: 38132.4 milliseconds
Iteration #7	 Margin per band per cell: 2.87123e+08
Time taken by inner loop: 38400.7 milliseconds
Iteration #8	 Margin per band per cell: 2.08513e+08
Time taken by inner loop: 38119.6 milliseconds
Iteration #9	 Margin per band per cell: 1.33949e+08
Time taken by inner loop: 38469.1 milliseconds
Iteration #10	 Margin per band per cell: 6.66506e+07
Time taken by inner loop: 38032.8 milliseconds
Iteration #11	 Margin per band per cell: 2.41749e+07
Time taken by inner loop: 38582.8 milliseconds
Iteration #12	 Margin per band per cell: 6.57021e+06
Time taken by inner loop: 38281 milliseconds
Iteration #13	 Margin per band per cell: 1.46716e+06
Time taken by inner loop: 38621.5 milliseconds
Iteration #14	 Margin per band per cell: 317888
Time taken by inner loop: 38287.2 milliseconds
Iteration #15	 Margin per band per cell: 84860.6
Time taken by inner loop: 39297.2 milliseconds
Iteration #16	 Margin per band per cell: 31922.9
Time taken by inner loop: 38627.1 milliseconds
Iteration #17	 Margin per band per cell: 15188.9
Time taken by inner loop: 38851.2 milliseconds
Iteration #18	 Margin per band per cell: 7789.49
Time taken by inner loop: 38479.7 milliseconds
Iteration #19	 Margin per band per cell: 3912.83
Time taken by inner loop: 39178.2 milliseconds
Iteration #20	 Margin per band per cell: 1832.17
Time taken by inner loop: 38715.1 milliseconds
Iteration #21	 Margin per band per cell: 780.45
Time taken by inner loop: 38853.1 milliseconds
Iteration #22	 Margin per band per cell: 300.722
Time taken by inner loop: 38356.2 milliseconds
Iteration #23	 Margin per band per cell: 106.085
Time taken by inner loop: 38911.4 milliseconds
Iteration #24	 Margin per band per cell: 35.3185
Time taken by inner loop: 38557.2 milliseconds
Iteration #25	 Margin per band per cell: 11.631
Time taken by inner loop: 38941.5 milliseconds
Iteration #26	 Margin per band per cell: 3.99096
Time taken by inner loop: 38529.1 milliseconds
Iteration #27	 Margin per band per cell: 1.4749
Time taken by inner loop: 38929.8 milliseconds
Iteration #28	 Margin per band per cell: 0.585538
Time taken by inner loop: 38698.5 milliseconds
Iteration #29	 Margin per band per cell: 0.242716
Time taken by inner loop: 39033.9 milliseconds
Iteration #30	 Margin per band per cell: 0.101684
Time taken by inner loop: 38623.6 milliseconds
Iteration #31	 Margin per band per cell: 0.0420221
Time taken by inner loop: 39082.4 milliseconds
Iteration #32	 Margin per band per cell: 0.0168996
Time taken by inner loop: 38765.3 milliseconds
Iteration #33	 Margin per band per cell: 0.0065871
Time taken by inner loop: 39099.1 milliseconds
Iteration #34	 Margin per band per cell: 0.00249834
Time taken by inner loop: 38680.5 milliseconds
Iteration #35	 Margin per band per cell: 0.00093141
Time taken by inner loop: 39374.5 milliseconds
Iteration #36	 Margin per band per cell: 0.00034594
Time taken by inner loop: 39058.5 milliseconds
Iteration #37	 Margin per band per cell: 0.000129646
Time taken by inner loop: 39313.7 milliseconds
Iteration #38	 Margin per band per cell: 4.9416e-05
Time taken by inner loop: 38921.2 milliseconds


Time taken by iteration: 1.50718e+06 milliseconds
heat flux of Boundary Condition #0: -1.29676e+09
heat flux of Boundary Condition #1: 619.333
heat flux of Boundary Condition #2: -1197.31
heat flux of Boundary Condition #3: 1.29678e+09
