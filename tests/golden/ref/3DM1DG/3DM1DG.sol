N_cell: 100
N_dir: 8
N_band: 1
num_theta: 8
num_phi: 1

Iteration #0	 Margin per band per cell: 9.85336e+15
Iteration #1	 Margin per band per cell: 1.79465e+14
Iteration #2	 Margin per band per cell: 2.75009e+12
Iteration #3	 Margin per band per cell: 4.18626e+10
Iteration #4	 Margin per band per cell: 6.37111e+08
Iteration #5	 Margin per band per cell: 9.69618e+06
Iteration #6	 Margin per band per cell: 147566
Iteration #7	 Margin per band per cell: 2245.81
Iteration #8	 Margin per band per cell: 34.1789
Iteration #9	 Margin per band per cell: 0.520168
Iteration #10	 Margin per band per cell: 0.00791643
Iteration #11	 Margin per band per cell: 0.00012048
Iteration #12	 Margin per band per cell: 1.83363e-06

heat flux of Boundary Condition #0: -2.0214e+10
heat flux of Boundary Condition #1: 2.0214e+10
