1 process(es),   1 device(s)
Bind solver rank 0 to device 0.
N_cell: 78190
N_dir: 32
N_band: 1
num_theta: 8
num_phi: 8

Iteration #0	 Margin per band per cell: 1.19597e+11
Time taken by inner loop: 3578.48 milliseconds
Iteration #1	 Margin per band per cell: 6.29534e+10
Time taken by inner loop: 4941.55 milliseconds
Iteration #2	 Margin per band per cell: 1.29824e+10
Time taken by inner loop: 4839.95 milliseconds
Iteration #3	 Margin per band per cell: 5.02185e+09
Time taken by inner loop: 4782.92 milliseconds
Iteration #4	 Margin per band per cell: 1.08479e+09
Time taken by inner loop: 4768.22 milliseconds
Iteration #5	 Margin per band per cell: 6.76156e+08
Time taken by inner loop: 4791.18 milliseconds
Iteration #6	 Margin per band per cell: 5.2108e+08
Time taken by inner loop: 4778.61 milliseconds
Iteration #7	 Margin per band per cell: 4.01901e+08
Time taken by inner loop: 4733.94 milliseconds
Iteration #8	 Margin per band per cell: 3.09301e+08
Time taken by inner loop: 4743.95 milliseconds
Iteration #9	 Margin per band per cell: 2.10663e+08
Time taken by inner loop: 4747.44 milliseconds
Iteration #10	 Margin per band per cell: 9.87691e+07
Time taken by inner loop: 4760.94 milliseconds
Iteration #11	 Margin per band per cell: 2.35772e+07
Time taken by inner loop: 4737.56 milliseconds
Iteration #12	 Margin per band per cell: 2.88604e+06
Time taken by inner loop: 4749.24 milliseconds
Iteration #13	 Margin per band per cell: 330314
Time taken by inner loop: 4738.52 milliseconds
Iteration #14	 Margin per band per cell: 85545.7
Time taken by inner loop: 4764.3 milliseconds
Iteration #15	 Margin per band per cell: 39389.6
Time taken by inner loop: 4758.31 milliseconds
Iteration #16	 Margin per band per cell: 22012.5
Time taken by inner loop: 4761.06 milliseconds
Iteration #17	 Margin per band per cell: 12842.4
Time taken by inner loop: 4762.86 milliseconds
Iteration #18	 Margin per band per cell: 7186.66
Time taken by inner loop: 4746.02 milliseconds
Iteration #19	 Margin per band per cell: 3619.1
Time taken by inner loop: 4782.55 milliseconds
Iteration #20	 Margin per band per cell: 1544.62
Time taken by inner loop: 4761.82 milliseconds
Iteration #21	 Margin per band per cell: 531.971
Time taken by inner loop: 4816.66 milliseconds
Iteration #22	 Margin per band per cell: 147.133
Time taken by inner loop: 4750.89 milliseconds
Iteration #23	 Margin per band per cell: 35.9155
Time taken by inner loop: 4763.18 milliseconds
Iteration #24	 Margin per band per cell: 9.57539
Time taken by inner loop: 4764.25 milliseconds
Iteration #25	 Margin per band per cell: 3.27499
Time taken by inner loop: 4750.3 milliseconds
Iteration #26	 Margin per band per cell: 1.36287
Time taken by inner loop: 4755.07 milliseconds
Iteration #27	 Margin per band per cell: 0.614034
Time taken by inner loop: 4731.82 milliseconds
Iteration #28	 Margin per band per cell: 0.277385
Time taken by inner loop: 4748.39 milliseconds
Iteration #29	 Margin per band per cell: 0.119948
Time taken by inner loop: 4775.36 milliseconds
Iteration #30	 Margin per band per cell: 0.0481939
Time taken by inner loop: 4764.07 milliseconds
Iteration #31	 Margin per band per cell: 0.017724
Time taken by inner loop: 4760.16 milliseconds
Iteration #32	 Margin per band per cell: 0.00599983
Time taken by inner loop: 4752.64 milliseconds
Iteration #33	 Margin per band per cell: 0.00193118
Time taken by inner loop: 4763.07 milliseconds
Iteration #34	 Margin per band per cell: 0.00062514
Time taken by inner loop: 4763.29 milliseconds
Iteration #35	 Margin per band per cell: 0.000215835
Time taken by inner loop: 4754.31 milliseconds
Iteration #36	 Margin per band per cell: 8.10436e-05
Time taken by inner loop: 4776.89 milliseconds


Time taken by iteration: 175222 milliseconds
heat flux of Boundary Condition #0: -1.29714e+09
heat flux of Boundary Condition #1: 88.0969
heat flux of Boundary Condition #2: -129.495
heat flux of Boundary Condition #3: 1.29713e+09
Time taken by solver: 564766 milliseconds
